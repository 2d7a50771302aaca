// refine.hip -- one synchronous greedy round on the GPU (refine/greedy.rs:36-193).
//
// A round of the reference visits the boundary vertices B in ascending order.
// Gains use the round-start partition, but feasibility (and the repair pass's
// "source overweight" test) use a running part-weight vector pw that is
// updated after every recorded move; the moves are applied after the scan.
// The GPU reproduces that bit-exactly in two launches:
//
//  1. k_gain_tile (64-vertex tiles, a wave each) / k_gain_heavy (a block per vertex of
//     degree > kGainHeavy): boundary test and the neighbour-part gain table
//     ed[p] = sum of edge weights to part p (compute_move_gain,
//     greedy.rs:197-218, gives gain(v,pv,p) = ed[p] - ed[pv]).  Every vertex
//     that could move gets a record and its candidate parts in admission
//     order -- descending gain, ties to the lowest part -- which is the order
//     in which the reference's strict `gain > best_gain` scan over p = 0..k-1
//     picks the first feasible part.  Parts with ed == 0 (all non-adjacent
//     parts) share one gain value and are kept implicit ("zero group").
//     Records are compacted per tile in vertex order.
//  2. admission (admit.hip): one workgroup replays the ordered scan.
#include <climits>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "profiler.hpp"

namespace rk {

namespace {

constexpr int kHeavyLds = 4096;  // k kept in LDS by the heavy gain kernel

struct GainArgs {
    int32_t n, ncon, k, mode;
    const int32_t* xadj;
    const int32_t* adjncy;
    const int32_t* adjwgt;
    const int32_t* vwgt;
    const int32_t* part;
    const int32_t* pw;          // round-start part weights
    const long long* omax;      // overweight thresholds
    const int4* heavy_rec;      // records of vertices handled by k_gain_heavy
    int32_t* pool;              // candidate parts, vertex v at pool[xadj[v]...]
    int4* tile_rec;
    int32_t* tile_cnt;
    uint32_t* tile_bits;
    const int32_t* done;
    int4* vrec;                 // optional per-vertex record (gain-table export)
    int32_t* g0_out;            // optional zero-group gain per vertex
    int32_t* gains;             // optional candidate gains beside pool
    // batched levels (null for a single graph): per tile its first vertex
    // (-1: padding), vertex count and graph; per vertex its graph; per graph
    // the offset of its k*ncon tables (pw, omax), its part count and its flag
    // "a cut pass moved nothing"
    const int32_t* tile_v0;
    const int32_t* tile_nv;
    const int32_t* tile_g;
    const int32_t* gseg;
    const int32_t* gkc;
    const int32_t* gk;
    const int32_t* gdone;
    // incremental cut pass: only tiles marked in tile_dirty (or every tile
    // of a graph whose full flag is set: full[g], or full[0] for one graph)
    // are recomputed; the others keep last pass's records
    uint32_t* tile_dirty;
    const int32_t* full;
    int incremental;
};

__device__ inline bool gain_qualifies(int mode, int32_t g) {
    // cut-optimisation: best_gain starts at 0 with strict > (greedy.rs:135,150);
    // repair: best_gain starts at i32::MIN with strict > (greedy.rs:78,93).
    return mode == 1 ? g > 0 : g != INT32_MIN;
}

__device__ inline bool part_overweight(const int32_t* pw, const long long* omax, int ncon, int32_t p) {
    for (int c = 0; c < ncon; ++c)
        if ((long long)pw[(int64_t)p * ncon + c] > omax[(int64_t)p * ncon + c]) return true;
    return false;
}

constexpr int kGWaves = 4;               // waves (independent tiles) per gain block
constexpr int kGCap = 1024;              // edges a wave stages in LDS per batch
constexpr int kGItems = kGCap / 64;      // staged edges per lane per batch
constexpr int kGainBlocksPerCU = 4;      // resident gain blocks per CU (LDS-limited)
static_assert(kTile == 64, "a gain tile is one wave");

__device__ inline int wave_excl_scan_i(int x, int* total) {
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

// Gain tables of one pass.  Every wave handles tiles of 64 consecutive
// vertices (lane = vertex) on its own, with its own LDS staging area and no
// block barriers, so the waves of a CU overlap their load and compute phases
// freely.  Waves are persistent over a strided sequence of tiles; the next
// tile's CSR offsets and own parts are loaded while the current one is
// processed.  Per tile:
//  * the edges of vertices with degree <= kGainHeavy are staged in batches
//    of whole vertices of at most kGCap edges: kGItems independent coalesced
//    loads per lane (adjncy, adjwgt), then the part[] gathers, into LDS;
//  * each lane builds its vertex's gain table from LDS: the weight to its
//    own part, the boundary test, and one entry per adjacent part with the
//    weight sum (aggregated in place over its edge range), ranked by
//    (gain desc, part asc) into pool[xadj[v]...];
//  * the tile's mover records are compacted with a ballot into
//    tile_rec[tile*64...], their count into tile_cnt[tile], and its bit in
//    tile_bits is set or cleared (the bitmap always describes the tiles'
//    current records), its bit in tile_dirty cleared.
// In an incremental cut pass the waves skip the tiles none of whose vertices
// or neighbours moved in the last pass: their records, counts and bits are
// still exact (a cut-pass record depends on the parts of the vertex and its
// neighbours only).
// BATCH: the level is a union of graphs (tile_v0/tile_nv/tile_g describe the
// tiles); instantiated separately so the single-graph kernel carries none of
// the batch branches.
// (At most 128 VGPRs: four waves per SIMD, the residency the LDS allows.)
template <bool BATCH>
__global__ __launch_bounds__(64 * kGWaves) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_gain_tile(GainArgs a,
                                                                                                       int ntiles) {
    __shared__ int2 s_e_all[kGWaves][kGCap];        // staged edges: (part of the neighbour, weight)
    __shared__ int32_t s_loff_all[kGWaves][kTile + 1];
    __shared__ int32_t s_beg_all[kGWaves][kTile];
    const int wv = threadIdx.x >> 6, lane = __lane_id();
    int2* s_e = s_e_all[wv];
    int32_t* s_loff = s_loff_all[wv];
    int32_t* s_beg = s_beg_all[wv];
    // head of the next tile: its span and, per lane, xadj[v], xadj[v+1], part[v]
    int64_t hv0 = 0;
    int hnv = 0;
    int32_t hb = 0, he = 0, hp = -1;
    auto load_head = [&](int tl) {
        if constexpr (BATCH) {
            hv0 = a.tile_v0[tl];
            hnv = a.tile_nv[tl];
        } else {
            hv0 = (int64_t)tl * kTile;
            hnv = (int)min((int64_t)kTile, (int64_t)a.n - hv0);
        }
        hb = he = 0;
        hp = -1;
        if (lane < hnv) {
            const int64_t v = hv0 + lane;
            hb = a.xadj[v];
            he = a.xadj[v + 1];
            hp = a.part[v];
        }
    };
    // tiles this pass recomputes (all of them unless incremental)
    auto live = [&](int tl) __attribute__((always_inline)) -> bool {
        if (!a.incremental) return true;
        int32_t f = a.full[0];
        if constexpr (BATCH) {
            const int32_t gi = a.tile_g[tl];
            if (gi < 0 || a.gdone[gi]) return false;   // padding, or converged: records already cleared
            f = a.full[gi];
        }
        return f != 0 || ((a.tile_dirty[tl >> 5] >> (tl & 31)) & 1u);
    };
    // the wave's tiles are tb + i*ts, i < per (XCD-contiguous regions)
    const WaveTiles wt = wave_tiles(ntiles, kGWaves);
    const int tb = wt.base, ts = wt.stride, per = wt.count;
    // live flags are fetched 64 tiles at a time (one load per lane) into a
    // ballot mask; the next live tile is the next set bit
    int cbase = -64;                                         // first index of the fetched chunk
    uint64_t cmask = 0;                                      // live tiles of the chunk
    auto next_live = [&](int i) __attribute__((always_inline)) -> int {
        for (;;) {
            if (i >= per) return per;
            if (i >= cbase + 64) {
                cbase = i - (i % 64);
                const int li = cbase + lane;
                cmask = __ballot(li < per && live(tb + li * ts));
            }
            const uint64_t m = cmask & (~0ull << (i - cbase));
            if (m) return cbase + __builtin_ctzll(m);
            i = cbase + 64;
        }
    };
    int it = next_live(0);
    if (it < per) load_head(tb + it * ts);
    if (!BATCH && *a.done) return;   // read together with the first head
    for (; it < per;) {
        const int tile = tb + it * ts;
        const int nit = next_live(it + 1);
        const int nxt = nit < per ? tb + nit * ts : ntiles;
        const int64_t v = hv0 + lane;
        const bool own = lane < hnv;
        const int32_t beg = hb, end = he, pv = hp;
        // this tile's graph tables (a single graph: the arguments themselves)
        const int32_t* pwg = a.pw;
        const long long* omg = a.omax;
        if constexpr (BATCH) {
            const int32_t gi = a.tile_g[tile];
            if (gi < 0 || a.gdone[gi]) {       // padding tile, or its graph has converged
                if (lane == 0) {
                    a.tile_cnt[tile] = 0;
                    atomicAnd(&a.tile_bits[tile >> 5], ~(1u << (tile & 31)));
                }
                load_head(nxt < ntiles ? nxt : tile);
                it = nit;
                continue;
            }
            pwg += a.gkc[gi];
            omg += a.gkc[gi];
        }
        // the next tile's head (the current one again after the last tile:
        // an unconditional load keeps the prefetch out of any branch)
        load_head(nxt < ntiles ? nxt : tile);
        const int32_t deg = end - beg;
        const bool heavy = own && deg > kGainHeavy;
        const int sdeg = (own && !heavy) ? deg : 0;
        int total;
        const int loff = wave_excl_scan_i(sdeg, &total);
        s_loff[lane] = loff;
        s_beg[lane] = beg;
        if (lane == 63) s_loff[kTile] = total;
        const bool anyheavy = __ballot(heavy) != 0;
        const int32_t x0 = __builtin_amdgcn_readfirstlane(beg);   // lane 0 owns v0 < n
        // A balance-repair pass moves only vertices of overweight parts
        // (greedy.rs:66-72): when none of the tile's ordinary vertices is in
        // one, none of them gets a record and their edges are not read (a
        // heavy vertex's record comes from k_gain_heavy either way).
        const bool skip =
            a.mode == 0 && __ballot(own && !heavy && part_overweight(pwg, omg, a.ncon, pv)) == 0;
        wave_sync();
        int4 rec = make_int4(-1, 0, 0, 0);
        // batches of whole vertices [vs, ve) with at most kGCap staged edges
        for (int vs = skip ? kTile : 0; vs < kTile;) {
            const int P0 = __shfl(loff, vs, 64);
            const uint64_t fit = __ballot(lane >= vs && loff + sdeg - P0 <= kGCap);
            const int ve = 64 - __builtin_clzll(fit);            // fit is lanes [vs, ve), ve > vs
            const int P1 = __shfl(loff + sdeg, ve - 1, 64);
            // All loads of a batch are kept in flight together: neighbour ids
            // first and weights second, then the part[] gathers (which wait
            // for the ids only), then the LDS writes.  Without heavy vertices
            // the batch is the edge range [x0+P0, x0+P1): buffer loads over
            // exactly that range (one offset register, lanes past the end
            // read 0 without a memory access).  Items past the batch gather
            // part[0].
            if (P1 > P0) {
                if (!anyheavy) {
                    int32_t uu[kGItems], ww[kGItems], pu[kGItems];
                    const int p0 = __builtin_amdgcn_readfirstlane(P0);
                    const int ne = __builtin_amdgcn_readfirstlane(P1 - P0);
                    const int64_t e0 = (int64_t)x0 + p0;
                    const int32_t* wsrc = a.adjwgt ? a.adjwgt : a.adjncy;
                    const int lb4 = lane * 4;   // the lane's byte offset; item i is a resource 64 edges on
#pragma unroll
                    for (int i = 0; i < kGItems; ++i)
                        uu[i] = __builtin_amdgcn_raw_buffer_load_b32(
                            edge_rsrc(a.adjncy + e0 + i * 64, max(ne - i * 64, 0)), lb4, 0, 0);
#pragma unroll
                    for (int i = 0; i < kGItems; ++i)
                        ww[i] = __builtin_amdgcn_raw_buffer_load_b32(
                            edge_rsrc(wsrc + e0 + i * 64, max(ne - i * 64, 0)), lb4, 0, 0);
#pragma unroll
                    for (int i = 0; i < kGItems; ++i) pu[i] = a.part[uu[i]];
#pragma unroll
                    for (int i = 0; i < kGItems; ++i) {
                        const int p = P0 + lane + i * 64;
                        if (p < P1) s_e[p - P0] = make_int2(pu[i], a.adjwgt ? ww[i] : 1);
                    }
                } else {
                    // a tile with a heavy vertex (rare): item by item
#pragma unroll 1
                    for (int i = 0; i < kGItems; ++i) {
                        const int p = P0 + lane + i * 64;
                        if (p < P1) {
                            int lo = vs, hi = ve - 1;   // last vertex whose staged range starts at or before p
                            while (lo < hi) {
                                const int mid = (lo + hi + 1) >> 1;
                                if (s_loff[mid] <= p) lo = mid;
                                else hi = mid - 1;
                            }
                            const int32_t j = s_beg[lo] + (p - s_loff[lo]);
                            s_e[p - P0] = make_int2(a.part[a.adjncy[j]], a.adjwgt ? a.adjwgt[j] : 1);
                        }
                    }
                }
            }
            wave_sync();
            if (lane >= vs && lane < ve && sdeg > 0) {
                const int lb = loff - P0, le = lb + deg;
                // weight to the own part and "any external edge"; four
                // independent LDS reads per step (the ranges of the lanes
                // differ, so each read is a gather)
                uint32_t internal = 0;
                bool ext = false;
                int e = lb;
                for (; e + 4 <= le; e += 4) {
                    const int2 x0 = s_e[e], x1 = s_e[e + 1], x2 = s_e[e + 2], x3 = s_e[e + 3];
                    internal += (x0.x == pv ? (uint32_t)x0.y : 0u) + (x1.x == pv ? (uint32_t)x1.y : 0u) +
                                (x2.x == pv ? (uint32_t)x2.y : 0u) + (x3.x == pv ? (uint32_t)x3.y : 0u);
                    ext |= (x0.x != pv) | (x1.x != pv) | (x2.x != pv) | (x3.x != pv);
                }
                for (; e < le; ++e) {
                    const int2 x = s_e[e];
                    if (x.x == pv) internal += (uint32_t)x.y;
                    else ext = true;
                }
                if (ext && (a.mode == 1 || part_overweight(pwg, omg, a.ncon, pv))) {
                    const int32_t ed_own = (int32_t)internal;
                    const int32_t g0 = wsub(0, ed_own);
                    int ncand = 0, nbefore = 0;
                    // The distinct adjacent parts with their weight sums: up
                    // to four in registers in one pass over the edges (the
                    // usual case), ranked pairwise; more fall back to the
                    // aggregation in LDS.
                    int32_t q0 = -1, q1 = -1, q2 = -1, q3 = -1;
                    uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
                    int nq = 0;
                    for (int e = lb; e < le; ++e) {
                        const int2 x = s_e[e];
                        if (x.x == pv) continue;
                        const uint32_t w = (uint32_t)x.y;
                        if (x.x == q0) w0 += w;
                        else if (x.x == q1) w1 += w;
                        else if (x.x == q2) w2 += w;
                        else if (x.x == q3) w3 += w;
                        else {
                            if (nq == 0) q0 = x.x, w0 = w;
                            else if (nq == 1) q1 = x.x, w1 = w;
                            else if (nq == 2) q2 = x.x, w2 = w;
                            else if (nq == 3) q3 = x.x, w3 = w;
                            ++nq;
                        }
                    }
                    if (nq <= 4) {
                        // a slot is a candidate when its sum is nonzero and its
                        // gain qualifies (a zero sum belongs to the zero group)
                        int32_t g[4];
                        bool c[4];
                        const int32_t qs[4] = {q0, q1, q2, q3};
                        const uint32_t ws[4] = {w0, w1, w2, w3};
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            g[j] = wsub((int32_t)ws[j], ed_own);
                            c[j] = j < nq && ws[j] != 0 && gain_qualifies(a.mode, g[j]);
                            ncand += c[j];
                            nbefore += c[j] && g[j] > g0;
                        }
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            int rank = 0;
#pragma unroll
                            for (int i = 0; i < 4; ++i)
                                rank += i != j && c[i] && (g[i] > g[j] || (g[i] == g[j] && qs[i] < qs[j]));
                            if (c[j]) {
                                a.pool[beg + rank] = qs[j];
                                if (a.gains) a.gains[beg + rank] = g[j];
                            }
                        }
                    } else {
                        // aggregate equal parts onto their first occurrence
                        for (int e = lb; e < le; ++e) {
                            const int32_t q = s_e[e].x;
                            if (q < 0 || q == pv) continue;
                            uint32_t sum = (uint32_t)s_e[e].y;
                            for (int e2 = e + 1; e2 < le; ++e2)
                                if (s_e[e2].x == q) {
                                    sum += (uint32_t)s_e[e2].y;
                                    s_e[e2].x = -1;
                                }
                            s_e[e].y = (int32_t)sum;
                        }
                        for (int e = lb; e < le; ++e) {
                            const int2 x = s_e[e];
                            const int32_t q = x.x;
                            if (q < 0 || q == pv || x.y == 0) continue;
                            const int32_t gq = wsub(x.y, ed_own);
                            if (!gain_qualifies(a.mode, gq)) continue;
                            ++ncand;
                            if (gq > g0) ++nbefore;
                            int rank = 0;
                            for (int e2 = lb; e2 < le; ++e2) {
                                const int2 x2 = s_e[e2];
                                const int32_t q2 = x2.x;
                                if (q2 < 0 || q2 == pv || x2.y == 0) continue;
                                const int32_t g2 = wsub(x2.y, ed_own);
                                if (!gain_qualifies(a.mode, g2)) continue;
                                if (g2 > gq || (g2 == gq && q2 < q)) ++rank;
                            }
                            a.pool[beg + rank] = q;
                            if (a.gains) a.gains[beg + rank] = gq;
                        }
                    }
                    const bool zflag = gain_qualifies(a.mode, g0);
                    if (a.g0_out) a.g0_out[v] = g0;
                    if (ncand > 0 || zflag)
                        rec = make_int4((int)v, pv, ncand, (int)((uint32_t)nbefore | (zflag ? 0x80000000u : 0u)));
                }
            }
            wave_sync();
            vs = ve;
        }
        if (heavy) rec = a.heavy_rec[v];
        if (a.vrec && own) a.vrec[v] = rec;
        const bool f = rec.x >= 0;
        const uint64_t fm = __ballot(f);
        if (f) a.tile_rec[(int64_t)tile * kTile + __popcll(fm & ((1ull << lane) - 1ull))] = rec;
        if (lane == 0) {
            const uint32_t bit = 1u << (tile & 31);
            a.tile_cnt[tile] = __popcll(fm);
            if (fm) atomicOr(&a.tile_bits[tile >> 5], bit);
            else atomicAnd(&a.tile_bits[tile >> 5], ~bit);
            if (a.tile_dirty) atomicAnd(&a.tile_dirty[tile >> 5], ~bit);
        }
        it = nit;
    }
}

// One block per heavy vertex; dense neighbour-part histogram of k entries in
// LDS (or in a per-block global scratch row when k > kHeavyLds).
__global__ __launch_bounds__(kBlock) void k_gain_heavy(GainArgs a, const int32_t* __restrict__ heavy_list,
                                                       int32_t nheavy, int4* __restrict__ heavy_rec,
                                                       int32_t* __restrict__ gscratch) {
    extern __shared__ int32_t s_dyn[];
    __shared__ uint32_t s_int;
    __shared__ int s_ext, s_nc, s_nb;
    if (!a.gseg && *a.done) return;
    const int kmax = a.k;   // the largest part count of the level's graphs
    int32_t* hist;
    int2* cands;
    if (kmax <= kHeavyLds) {
        hist = s_dyn;
        cands = (int2*)(s_dyn + kmax + (kmax & 1));
    } else {
        hist = gscratch + (int64_t)blockIdx.x * 3 * kmax;
        cands = (int2*)(hist + kmax + (kmax & 1));
    }
    for (int hi = blockIdx.x; hi < nheavy; hi += gridDim.x) {
        const int32_t v = heavy_list[hi];
        int k = kmax;
        const int32_t* pwg = a.pw;
        const long long* omg = a.omax;
        if (a.gseg) {
            const int32_t gi = a.gseg[v];
            if (a.gdone[gi]) {
                if (threadIdx.x == 0) heavy_rec[v] = make_int4(-1, 0, 0, 0);
                continue;                      // block-uniform
            }
            k = a.gk[gi];
            pwg += a.gkc[gi];
            omg += a.gkc[gi];
        }
        const int32_t beg = a.xadj[v], end = a.xadj[v + 1];
        const int32_t pv = a.part[v];
        for (int p = threadIdx.x; p < k; p += kBlock) hist[p] = 0;
        if (threadIdx.x == 0) {
            s_int = 0;
            s_ext = 0;
            s_nc = 0;
            s_nb = 0;
        }
        __syncthreads();
        uint32_t internal = 0;
        int ext = 0;
        for (int32_t j = beg + threadIdx.x; j < end; j += kBlock) {
            const int32_t u = a.adjncy[j];
            const int32_t w = a.adjwgt ? a.adjwgt[j] : 1;
            const int32_t pu = a.part[u];
            if (pu == pv) internal += (uint32_t)w;
            else {
                ext = 1;
                atomicAdd(&hist[pu], w);
            }
        }
        internal = group_sum<64>(internal);
        const bool wave_ext = __ballot(ext) != 0;
        if (__lane_id() == 0) {
            atomicAdd(&s_int, internal);
            if (wave_ext) atomicOr(&s_ext, 1);
        }
        __syncthreads();
        const int32_t ed_own = (int32_t)s_int;
        int4 rec = make_int4(-1, 0, 0, 0);
        const bool gate = a.mode == 1 || part_overweight(pwg, omg, a.ncon, pv);
        if (s_ext && gate) {
            const int32_t g0 = wsub(0, ed_own);
            for (int p = threadIdx.x; p < k; p += kBlock) {
                const int32_t val = hist[p];
                if (p == pv || val == 0) continue;
                const int32_t g = wsub(val, ed_own);
                if (!gain_qualifies(a.mode, g)) continue;
                const int slot = atomicAdd(&s_nc, 1);
                cands[slot] = make_int2(p, g);
                if (g > g0) atomicAdd(&s_nb, 1);
            }
            __syncthreads();
            const int nc = s_nc;
            for (int e = threadIdx.x; e < nc; e += kBlock) {
                const int2 me = cands[e];
                int rank = 0;
                for (int e2 = 0; e2 < nc; ++e2) {
                    const int2 o = cands[e2];
                    if (o.y > me.y || (o.y == me.y && o.x < me.x)) ++rank;
                }
                a.pool[beg + rank] = me.x;
                if (a.gains) a.gains[beg + rank] = me.y;
            }
            if (a.g0_out && threadIdx.x == 0) a.g0_out[v] = g0;
            const bool zflag = gain_qualifies(a.mode, g0);
            if (nc > 0 || zflag)
                rec = make_int4(v, pv, nc, (int)((uint32_t)s_nb | (zflag ? 0x80000000u : 0u)));
        }
        if (threadIdx.x == 0) heavy_rec[v] = rec;
        __syncthreads();
    }
}

// Tiles the next, incremental cut pass recomputes: those of the moved
// vertices and of their neighbours (one wave per move); a hub sets *full.
__global__ __launch_bounds__(kBlock) void k_mark_moves(const int2* __restrict__ moves, const int32_t* __restrict__ nmoves,
                                                       const int32_t* __restrict__ xadj,
                                                       const int32_t* __restrict__ adjncy, uint32_t* __restrict__ dirty,
                                                       int32_t* __restrict__ full, int32_t max_deg) {
    const int32_t nm = *nmoves;
    const int lane = __lane_id();
    const int64_t nw = (int64_t)gridDim.x * (kBlock / 64);
    for (int64_t i = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); i < nm; i += nw) {
        const int32_t v = moves[i].x;
        const int32_t b0 = xadj[v], b1 = xadj[v + 1];
        if (lane == 0) {
            const int32_t t = v >> 6;
            atomicOr(&dirty[t >> 5], 1u << (t & 31));
            if (b1 - b0 > max_deg) *full = 1;
        }
        if (b1 - b0 > max_deg) continue;
        for (int32_t j = b0 + lane; j < b1; j += 64) {
            const int32_t t = adjncy[j] >> 6;
            atomicOr(&dirty[t >> 5], 1u << (t & 31));
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_heavy_list(int32_t n_max, const int32_t* __restrict__ n_dev,
                                                       const int32_t* __restrict__ xadj, int32_t thresh,
                                                       int32_t* __restrict__ list, int32_t* __restrict__ count) {
    const int32_t n = n_dev ? *n_dev : n_max;
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool h = v < n && (xadj[v + 1] - xadj[v]) > thresh;
    uint64_t b = __ballot(h);
    int base = 0;
    if (__lane_id() == 0 && b) base = atomicAdd(count, __popcll(b));
    base = __shfl(base, 0, 64);
    if (h) list[base + __popcll(b & ((1ull << __lane_id()) - 1ull))] = (int32_t)v;
}

void launch_gain(const GainArgs& ga, int nheavy, const int32_t* heavy_list, RefineBufs& b, hipStream_t s,
                 int ntiles = -1) {
    if (ntiles < 0) ntiles = (ga.n + kTile - 1) / kTile;
    const int k = ga.k;
    if (nheavy > 0) {
        const size_t lds = k <= kHeavyLds ? (size_t)(k + (k & 1) + 2 * k) * sizeof(int32_t) : 0;
        const int cap = k <= kHeavyLds ? 2048 : kHeavyGlobalBlocks;
        const int grid = nheavy < cap ? nheavy : cap;
        hipLaunchKernelGGL(k_gain_heavy, dim3(grid), dim3(kBlock), lds, s, ga, heavy_list, nheavy, b.heavy_rec,
                           b.heavy_scratch);
    }
    const int blocks = (ntiles + kGWaves - 1) / kGWaves, resident = device_cus() * kGainBlocksPerCU;
    const dim3 grid(blocks < resident ? blocks : resident);
    if (ga.tile_g) hipLaunchKernelGGL(k_gain_tile<true>, grid, dim3(64 * kGWaves), 0, s, ga, ntiles);
    else hipLaunchKernelGGL(k_gain_tile<false>, grid, dim3(64 * kGWaves), 0, s, ga, ntiles);
}

}  // namespace

void gain_table(const DevGraph& g, int32_t k, int nheavy, const int32_t* heavy_list, RefineBufs& b,
                const int32_t* part, int4* vrec, int32_t* g0, int32_t* gains, hipStream_t s) {
    if (g.n <= 0) return;
    // Every part overweight (pw 0 > omax = INT64_MIN) so that every boundary
    // vertex lists all its adjacent parts with a nonzero edge-weight sum.
    const size_t kc = (size_t)k * g.ncon;
    RK_HIP(hipMemsetAsync(b.pw, 0, kc * sizeof(int32_t), s));
    std::vector<long long> om(kc, LLONG_MIN);
    RK_HIP(hipMemcpyAsync(b.omax, om.data(), kc * sizeof(long long), hipMemcpyHostToDevice, s));
    RK_HIP(hipMemsetAsync(b.done, 0, sizeof(int32_t), s));
    GainArgs ga{g.n, g.ncon, k, 0, g.xadj, g.adjncy, g.adjwgt, g.vwgt, part, b.pw, b.omax,
                b.heavy_rec, b.pool, b.tile_rec, b.tile_cnt, b.tile_bits, b.done, vrec, g0, gains};
    launch_gain(ga, nheavy, heavy_list, b, s);
    RK_HIP(hipStreamSynchronize(s));
    RK_HIP(hipGetLastError());
}

void fm_candidates(const DevGraph& g, int32_t k, int nheavy, const int32_t* heavy_list, RefineBufs& b,
                   const int32_t* part, const int32_t* pw0, const long long* omin, const int32_t* done, int4* vrec,
                   int32_t* g0, int32_t* gains, hipStream_t s) {
    if (g.n <= 0) return;
    GainArgs ga{g.n, g.ncon, k, 0, g.xadj, g.adjncy, g.adjwgt, g.vwgt, part, pw0, omin,
                b.heavy_rec, b.pool, b.tile_rec, b.tile_cnt, b.tile_bits, done, vrec, g0, gains};
    launch_gain(ga, nheavy, heavy_list, b, s);
    RK_HIP(hipGetLastError());
}

void heavy_list_build(int32_t n_max, const int32_t* n_dev, const int32_t* xadj, int32_t* list, int32_t* count,
                      hipStream_t s, bool count_zeroed) {
    if (!count_zeroed) RK_HIP(hipMemsetAsync(count, 0, sizeof(int32_t), s));
    if (n_max > 0)
        hipLaunchKernelGGL(k_heavy_list, dim3((n_max + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n_max, n_dev,
                           xadj, kGainHeavy, list, count);
    RK_HIP(hipGetLastError());
}

int heavy_list_sync(const DevGraph& g, int32_t* list, RefineBufs& b, hipStream_t s) {
    heavy_list_build(g.n, nullptr, g.xadj, list, b.heavy_cnt, s);
    RK_HIP(hipMemcpyAsync(b.host_small, b.heavy_cnt, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RK_HIP(hipStreamSynchronize(s));
    return b.host_small[0];
}

void refine_pass(const DevGraph& g, int32_t k, int mode, int nheavy, const int32_t* heavy_list, RefineBufs& b,
                 int32_t* part, hipStream_t s, bool incremental, bool mark_next) {
    if (g.n <= 0) return;
    const int ntiles = (g.n + kTile - 1) / kTile;
    GainArgs ga{g.n, g.ncon, k, mode, g.xadj, g.adjncy, g.adjwgt, g.vwgt, part, b.pw, b.omax,
                b.heavy_rec, b.pool, b.tile_rec, b.tile_cnt, b.tile_bits, b.done, nullptr, nullptr, nullptr};
    ga.tile_dirty = b.tile_dirty;
    ga.full = b.full;
    ga.incremental = incremental && b.tile_dirty && b.full && mode == 1 ? 1 : 0;
    {
        // algorithmic bytes of a full launch: xadj, adjncy, adjwgt, part[adjncy] gather, part[v]
        const double bytes = 4.0 * (g.n + 1) + 4.0 * g.n + (g.adjwgt ? 12.0 : 8.0) * g.m;
        // an incremental launch, and a balance-repair launch (which reads
        // the rows of tiles with a vertex in an overweight part only), read a
        // data-dependent part of the level: their bytes are not modelled
        const bool partial = ga.incremental || mode == 0;
        ProfScope ps(partial ? kProfGainInc : kProfGain, partial ? 0.0 : bytes, g.n, s);
        launch_gain(ga, nheavy, heavy_list, b, s);
    }
    {
        ProfScope ps(kProfAdmit, 0.0, g.n, s);
        admit_pass(g, k, mode, ntiles, b, part, s);
    }
    if (mark_next && mode == 1 && b.tile_dirty && b.full)   // one wave per move, not on the admission's chain
        hipLaunchKernelGGL(k_mark_moves, dim3(256), dim3(kBlock), 0, s, b.moves, b.nmoves, g.xadj, g.adjncy,
                           b.tile_dirty, b.full, kMarkMaxDeg);
    RK_HIP(hipGetLastError());
}

void refine_pass_batch(const DevGraph& g, const BatchLevel& bl, int mode, int nheavy, const int32_t* heavy_list,
                       RefineBufs& b, int32_t* part, hipStream_t s, bool incremental) {
    if (g.n <= 0 || bl.ngraphs <= 0) return;
    GainArgs ga{g.n, g.ncon, bl.kmax, mode, g.xadj, g.adjncy, g.adjwgt, g.vwgt, part, b.pw, b.omax,
                b.heavy_rec, b.pool, b.tile_rec, b.tile_cnt, b.tile_bits, b.done, nullptr, nullptr, nullptr,
                bl.tile_v0, bl.tile_nv, bl.tile_g, bl.gseg, bl.gkc, bl.gk, bl.gdone};
    ga.tile_dirty = b.tile_dirty;
    ga.full = bl.gfull;
    ga.incremental = incremental && b.tile_dirty && bl.gfull && mode == 1 ? 1 : 0;
    {
        const double bytes = 4.0 * (g.n + 1) + 4.0 * g.n + (g.adjwgt ? 12.0 : 8.0) * g.m;
        // an incremental launch, and a balance-repair launch (which reads
        // the rows of tiles with a vertex in an overweight part only), read a
        // data-dependent part of the level: their bytes are not modelled
        const bool partial = ga.incremental || mode == 0;
        ProfScope ps(partial ? kProfGainInc : kProfGain, partial ? 0.0 : bytes, g.n, s);
        launch_gain(ga, nheavy, heavy_list, b, s, bl.ntiles);
    }
    {
        ProfScope ps(kProfAdmit, 0.0, g.n, s);
        admit_pass_batch(mode, bl.ngraphs, bl.admit, bl.kmax, g.ncon, s);
    }
    RK_HIP(hipGetLastError());
}

}  // namespace rk
