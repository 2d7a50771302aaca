// fm2.hip -- 2-way FM refinement (refine/fm.rs:12-134) for every graph of a
// batched level, one workgroup per graph.
//
// A pass of the reference computes the gain of every vertex (fm.rs:44-49),
// queues the boundary vertices by gain (52-57), then pops the best vertex
// until the queue is empty: a move that keeps the other side within its
// limit is made and locked, the cut tracked, and every unlocked neighbour's
// gain recomputed (requeued, or queued when it is now on the boundary); a
// move that does not fit is dropped (66-119).  Finally the moves after the
// best cut seen are undone (121-131); the pass "improved" when that best cut
// is below the cut of the partition before the undo, and FM stops after the
// first pass that did not improve (fm.rs:22-26).
//
// The pops are inherently ordered (each reads the part weights and gains the
// earlier ones left), so one wave runs a graph's pops while the graphs of
// the level run side by side:
//  * k_fm2_init (grid-wide): every vertex's gain and boundary flag, its
//    queue key and its state byte (side, locked, queued);
//  * k_fm2_tree (grid-wide, one launch per tree level): a 64-ary max tree
//    of keys above each graph's leaves;
//  * k_fm2_pass: the workgroup copies the graph's queue and vertex state
//    into LDS when they fit (else the tree and state bytes alone, else
//    nothing), then its
//    first wave runs the pops without block
//    barriers: every lane reads the root, lane 0 decides the move, the lanes
//    recompute the moved vertex's unlocked neighbours from their rows (one
//    neighbour each) and update the tree nodes above the changed leaves,
//    one level per wave barrier;
//  * the cut of the partition after the pops comes from the batched edge
//    cut, then k_fm2_undo undoes the moves past the best prefix.
// A queue key is (gain, then the smaller vertex first), the order among
// equal gains that this build fixes (the reference's queue, priority-queue
// 1.4.0, leaves it to that crate's heap).
#include <climits>

#include "common.hpp"
#include "kernels.hpp"

namespace rk {

namespace {

constexpr int kF2Threads = 256;
constexpr int kF2MaxCon = 16;      // constraints held in LDS by the pops
constexpr int kF2G = 8;            // lanes per vertex in k_fm2_init
constexpr int kF2Fan = 64;         // children per tree node (one wave)
constexpr int kF2FanLog = 6;
constexpr size_t kF2Lds = 156 * 1024;   // dynamic LDS of a pass: leaves, tree and state bytes

constexpr uint8_t kSide = 1, kLocked = 2, kQueued = 4;

struct Fm2Args {
    int32_t ncon;
    const int32_t* xadj;
    const int32_t* adjncy;
    const int32_t* adjwgt;
    const int32_t* vwgt;
    int32_t* part;
    int32_t* pw;                // per graph at gkc[g]: [2][ncon]
    const long long* fmax;      // per graph at gkc[g]: floor(target * ub)
    const int32_t* gkc;
    const int32_t* gseg;        // graph of each vertex
    const int32_t* gdesc_of;    // index into graphs[] of each graph present (-1: absent)
    const Fm2Graph* graphs;     // the graphs present
    int32_t ngraphs;
    uint64_t* leaf;             // [n] queue keys (0: not queued)
    uint64_t* tree;             // tree nodes above the leaves, a region per graph
    uint8_t* state;             // [n] side | locked | queued
    int32_t* gain;              // [n] gain of moving the vertex to the other side
    int32_t* ext;               // [n] entries of its row on the other side
    const int32_t* sym;         // 1: every row entry (v, u, w) has its mirror (u, v, w) (pair-symmetric)
    int32_t* moves;             // [n] the graph's moves at its first vertex
    const int32_t* cut0;        // per graph: edge cut at the pass start
    const int32_t* cut1;        // per graph: edge cut after the pops
    int32_t* best_len;          // per graph
    int32_t* best_cut;
    int32_t* nmoves;
    int32_t* done;              // per graph: FM has stopped on this level
    int32_t* err;
};

__device__ inline uint64_t f2_key(int32_t gain, int32_t lv) {
    return ((uint64_t)((uint32_t)gain ^ 0x80000000u) << 32) | (uint32_t)(0x7FFFFFFF - lv);
}

// Node count at level l of the tree of a graph of n vertices (level 0 =
// the leaves) and the offset of level l >= 1 in the graph's tree region.
// Computed by loops: arrays indexed by a run-time level would live in
// scratch memory.
__device__ inline int64_t lvl_len(int32_t n, int l) {
    int64_t len = n;
    for (int j = 0; j < l; ++j) len = (len + kF2Fan - 1) / kF2Fan;
    return len;
}
__device__ inline int64_t lvl_off(int32_t n, int l) {
    int64_t len = n, off = 0;
    for (int j = 1; j < l; ++j) {
        len = (len + kF2Fan - 1) / kF2Fan;
        off += len;
    }
    return off;
}
__device__ inline int tree_levels(int32_t n) {
    int l = 1;
    for (int64_t len = n; len > 1; len = (len + kF2Fan - 1) / kF2Fan) ++l;
    return l;
}

__device__ inline void wave_fence_sync() {
    // the lanes' earlier writes are visible to the whole wave: a wave's
    // memory operations are performed in order, so wavefront scope needs no
    // wait, only that the compiler keeps the order
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Gains, boundary and queue leaves of every vertex of the graphs still
// refining (fm.rs:44-57), and the state bytes.
__global__ __launch_bounds__(kBlock) void k_fm2_init(Fm2Args a, int32_t n) {
    const int gl = group_lane<kF2G>();
    for (int64_t v = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kF2G; v < n;
         v += (int64_t)gridDim.x * (kBlock / kF2G)) {
        const int32_t g = a.gseg[v];
        const int32_t di = a.gdesc_of[g];
        if (di < 0 || a.done[g]) continue;                       // group-uniform
        const int32_t pv = a.part[v];
        int32_t gain = 0, next = 0;
        for (int32_t j = a.xadj[v] + gl; j < a.xadj[v + 1]; j += kF2G) {
            const int32_t w = a.adjwgt ? a.adjwgt[j] : 1;
            if (a.part[a.adjncy[j]] == pv) {
                gain = wsub(gain, w);
            } else {
                gain = wadd(gain, w);
                ++next;
            }
        }
        gain = group_sum<kF2G>(gain);
        next = group_sum<kF2G>(next);
        if (gl == 0) {
            const int32_t lv = (int32_t)(v - a.graphs[di].v0);
            a.leaf[v] = next > 0 ? f2_key(gain, lv) : 0ull;
            a.state[v] = (uint8_t)((pv & 1) | (next > 0 ? kQueued : 0));
            a.gain[v] = gain;
            a.ext[v] = next;
        }
    }
}

// Level l (>= 1) of every graph's tree: one thread per node.  nodes[i] =
// the prefix over the graphs present of their node counts at level l.
__global__ __launch_bounds__(kBlock) void k_fm2_tree(Fm2Args a, int l, const int64_t* __restrict__ nodes) {
    const int64_t total = nodes[a.ngraphs];
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
        int lo = 0, hi = a.ngraphs - 1;   // the graph holding node t
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (nodes[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        const Fm2Graph gd = a.graphs[lo];
        if (a.done[gd.g]) continue;
        const int64_t i = t - nodes[lo];
        const uint64_t* src = l == 1 ? a.leaf + gd.v0 : a.tree + gd.tbase + lvl_off(gd.n, l - 1);
        uint64_t x = 0;
        const int64_t c0 = i * kF2Fan, c1 = min(c0 + kF2Fan, lvl_len(gd.n, l - 1));
        for (int64_t c = c0; c < c1; ++c) {
            const uint64_t y = src[c];
            x = y > x ? y : x;
        }
        a.tree[gd.tbase + lvl_off(gd.n, l) + i] = x;
    }
}

// One max step of a 64-bit value with the lane that the DPP pattern CTRL
// names, in the rows ROWS (lanes whose source is outside the row keep x).
template <int CTRL, int ROWS>
__device__ inline uint64_t dpp_max_step(uint64_t x) {
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    const uint32_t ylo = (uint32_t)__builtin_amdgcn_update_dpp((int)lo, (int)lo, CTRL, ROWS, 0xF, false);
    const uint32_t yhi = (uint32_t)__builtin_amdgcn_update_dpp((int)hi, (int)hi, CTRL, ROWS, 0xF, false);
    const uint64_t y = ((uint64_t)yhi << 32) | ylo;
    return y > x ? y : x;
}

// 64-bit maximum over the wave (every lane active) by DPP row shifts and
// row broadcasts; lane 63's result, as a wave-uniform value.
__device__ inline uint64_t wave_max_dpp(uint64_t x) {
    x = dpp_max_step<0x111, 0xF>(x);   // row_shr:1
    x = dpp_max_step<0x112, 0xF>(x);   // row_shr:2
    x = dpp_max_step<0x114, 0xF>(x);   // row_shr:4
    x = dpp_max_step<0x118, 0xF>(x);   // row_shr:8: lane 15 of each row holds the row's maximum
    x = dpp_max_step<0x142, 0xA>(x);   // row_bcast:15 into rows 1 and 3
    x = dpp_max_step<0x143, 0xC>(x);   // row_bcast:31 into rows 2 and 3: lane 63 holds the maximum
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), 63);
    return ((uint64_t)hi << 32) | lo;
}

// The tree above the changed leaves x0 and x1 of each lane (-1: none), one
// level per wave barrier; every lane of the wave takes part.  Level l >= 1
// sits at tree + s_off[l] with s_len[l] nodes (LDS tables).  Per level each
// distinct parent node of the changed entries (changed leaves are mostly
// neighbours, which share parents) is recomputed once: lane c reads child
// c, a DPP maximum over the wave, lane 0 writes.
__device__ inline void f2_tree_update(int nlvl, const int64_t* s_len, const int64_t* s_off, uint64_t* leaves,
                                      uint64_t* tree, int32_t x0, int32_t x1) {
    const int lane = __lane_id();
    for (int l = 1; l < nlvl; ++l) {
        const int sh = kF2FanLog * l;
        const int32_t na = x0 >= 0 ? (int32_t)(x0 >> sh) : -1;
        const int32_t nb = x1 >= 0 ? (int32_t)(x1 >> sh) : -1;
        const int64_t clen = s_len[l - 1];
        const uint64_t* src = l == 1 ? leaves : tree + s_off[l - 1];
        uint64_t* dst = tree + s_off[l];
        bool ra = na >= 0, rb = nb >= 0;
        for (;;) {
            const uint64_t left = __ballot(ra || rb);
            if (left == 0) break;
            const int f = __builtin_ctzll(left);
            const int32_t nd = __builtin_amdgcn_readlane(ra ? na : nb, f);
            ra = ra && na != nd;
            rb = rb && nb != nd;
            const int64_t child = (int64_t)nd * kF2Fan + lane;
            const uint64_t m = wave_max_dpp(child < clen ? src[child] : 0ull);
            if (lane == 0) dst[nd] = m;
        }
        wave_fence_sync();
    }
}

// The pops of one pass (fm.rs:65-119) for graph blockIdx.x.  MODE says
// where the graph's queue and vertex state live during the pops: 2 = all in
// LDS (tree, state bytes, gains and boundary counts with pair-symmetric
// rows, leaves), for graphs where they all fit; 1 = the tree above the
// leaves and the state bytes in LDS, leaves, gains and counts in global
// memory, for graphs where only those fit; 0 = all in global memory.  Each
// launch takes the graphs of its own mode and skips the others; separate
// instantiations keep every access of an LDS array an LDS instruction.
// Whether a graph of n vertices has everything of a pass in LDS even with
// gains and counts (pair-symmetric rows): then only the first pass launch
// can have work.
inline bool fm2_fits_lds(int32_t n) {
    const size_t tree_b = (size_t)fm2_tree_nodes(n) * 8, state_b = ((size_t)n + 15) & ~(size_t)15;
    return tree_b + state_b + (size_t)n * 16 <= kF2Lds;
}

template <int MODE>
__global__ __launch_bounds__(kF2Threads) void k_fm2_pass(Fm2Args a) {
    extern __shared__ __align__(16) unsigned char s_dyn[];
    __shared__ int32_t s_pw[2 * kF2MaxCon];
    __shared__ long long s_fm[2 * kF2MaxCon];
    __shared__ int64_t s_len[kFm2MaxLvl], s_off[kFm2MaxLvl];
    const Fm2Graph gd = a.graphs[blockIdx.x];
    if (a.done[gd.g]) return;
    const int tid = threadIdx.x, lane = __lane_id();
    const int ncon = a.ncon;
    const int32_t n = gd.n, v0 = gd.v0;
    // With pair-symmetric rows a move updates its neighbours' gains and
    // boundary counts from its own row (each entry (v, u, w) has the mirror
    // (u, v, w), so u's recomputed gain is the old one plus or minus 2w);
    // otherwise each neighbour's gain is recomputed from its row, as the
    // reference does.
    const bool sym = *a.sym != 0;
    const int nlvl = tree_levels(n);
    const int64_t nodes = lvl_off(n, nlvl);       // the tree region's size
    const size_t tree_b = (size_t)nodes * 8, state_b = ((size_t)n + 15) & ~(size_t)15;
    const size_t ge_b = sym ? (size_t)n * 8 : 0, leaf_b = (size_t)n * 8;
    const int mode = tree_b + state_b + ge_b + leaf_b <= kF2Lds ? 2 : (tree_b + state_b <= kF2Lds ? 1 : 0);
    if (mode != MODE) return;
    int32_t* pwg = a.pw + a.gkc[gd.g];
    const long long* fm = a.fmax + a.gkc[gd.g];
    if (tid < nlvl) {
        s_len[tid] = lvl_len(n, tid);
        s_off[tid] = lvl_off(n, tid);
    }
    uint64_t* tree;
    uint8_t* state;
    int32_t* gains;
    int32_t* exts;
    uint64_t* leaves;
    if constexpr (MODE == 2) {
        tree = (uint64_t*)s_dyn;
        state = s_dyn + tree_b;
        gains = (int32_t*)(s_dyn + tree_b + state_b);
        exts = gains + n;
        leaves = (uint64_t*)(s_dyn + tree_b + state_b + ge_b);
        for (int64_t i = tid; i < nodes; i += kF2Threads) tree[i] = a.tree[gd.tbase + i];
        for (int32_t i = tid; i < n; i += kF2Threads) {
            state[i] = a.state[v0 + i];
            leaves[i] = a.leaf[v0 + i];
        }
        if (sym)
            for (int32_t i = tid; i < n; i += kF2Threads) {
                gains[i] = a.gain[v0 + i];
                exts[i] = a.ext[v0 + i];
            }
    } else if constexpr (MODE == 1) {
        tree = (uint64_t*)s_dyn;
        state = s_dyn + tree_b;
        gains = a.gain + v0;
        exts = a.ext + v0;
        leaves = a.leaf + v0;
        for (int64_t i = tid; i < nodes; i += kF2Threads) tree[i] = a.tree[gd.tbase + i];
        for (int32_t i = tid; i < n; i += kF2Threads) state[i] = a.state[v0 + i];
    } else {
        tree = a.tree + gd.tbase;
        state = a.state + v0;
        gains = a.gain + v0;
        exts = a.ext + v0;
        leaves = a.leaf + v0;
    }
    for (int i = tid; i < 2 * ncon; i += kF2Threads) {
        s_pw[i] = pwg[i];
        s_fm[i] = fm[i];
    }
    {
        // Read the graph's rows once with the whole workgroup, so that the
        // pops' row reads (one wave, one vertex after another) find them in
        // the L2 cache instead of waiting on HBM each time.  The sum only
        // keeps the loads.
        const int32_t e0 = a.xadj[v0], e1 = a.xadj[v0 + n];
        uint32_t sum = 0;
        for (int64_t j = e0 + tid; j < e1; j += kF2Threads)
            sum += (uint32_t)a.adjncy[j] + (a.adjwgt ? (uint32_t)a.adjwgt[j] : 0u);
        for (int64_t j = v0 + tid; j <= (int64_t)v0 + n; j += kF2Threads) sum += (uint32_t)a.xadj[j];
        if (a.vwgt)
            for (int64_t j = (int64_t)v0 * ncon + tid; j < ((int64_t)v0 + n) * ncon; j += kF2Threads)
                sum += (uint32_t)a.vwgt[j];
        if (sum == 0x9e3779b9u) atomicOr(a.err, 0);   // changes nothing
    }
    __syncthreads();
    if (tid >= 64) return;                       // one wave runs the pops
    const uint64_t* root = nlvl == 1 ? leaves : tree + s_off[nlvl - 1];
    // every pop takes a queued vertex off the queue; a vertex is queued again
    // only when a neighbour moves, and each vertex moves at most once
    int64_t budget = (int64_t)n + (a.xadj[v0 + n] - a.xadj[v0]) + 1;
    int32_t cut = a.cut0[gd.g], best_cut = cut, best = 0, nm = 0;   // the same on every lane
    for (;;) {
        const uint64_t key = *root;
        if (key == 0) break;
        if (--budget < 0) {
            if (lane == 0) atomicExch(a.err, 1);
            break;
        }
        const int32_t i = 0x7FFFFFFF - (int32_t)(uint32_t)key;
        const int32_t gain = (int32_t)((uint32_t)(key >> 32) ^ 0x80000000u);
        const int32_t v = v0 + i;
        // v's row bounds are read together with its weights, before the
        // move is decided (they are needed only when it moves)
        const int32_t beg = a.xadj[v], deg = a.xadj[v + 1] - beg;
        const uint8_t sv = state[i];
        const int32_t from = sv & kSide, to = from ^ 1;
        int32_t vw = 0;
        bool fits = true;
        if (lane < ncon) {                                                    // fm.rs:74-81
            vw = a.vwgt ? a.vwgt[(int64_t)v * ncon + lane] : 1;
            fits = (long long)wadd(s_pw[to * ncon + lane], vw) <= s_fm[to * ncon + lane];
        }
        const bool feas = __ballot(!fits) == 0;
        wave_fence_sync();
        if (lane == 0) leaves[i] = 0;                                          // popped
        if (feas) {                                                           // fm.rs:83-97
            if (lane < ncon) {
                s_pw[from * ncon + lane] = wsub(s_pw[from * ncon + lane], vw);
                s_pw[to * ncon + lane] = wadd(s_pw[to * ncon + lane], vw);
            }
            if (lane == 0) {
                state[i] = (uint8_t)(((sv ^ kSide) | kLocked) & ~kQueued);
                a.moves[v0 + nm] = i;
            }
            ++nm;
            cut = wsub(cut, gain);
            if (cut < best_cut) {
                best_cut = cut;
                best = nm;
            }
        } else if (lane == 0) {
            state[i] = (uint8_t)(sv & ~kQueued);
        }
        wave_fence_sync();
        // the popped leaf's path is updated with the first chunk of
        // neighbours, or on its own (every lane takes part in an update)
        int32_t c0 = lane == 0 ? i : -1;
        bool pending = true;
        if (feas && sym) {
            // the unlocked neighbours' gains and boundary counts (fm.rs:99-118)
            // from v's row: an entry to v's old side turns external (+2w),
            // one to its new side internal (-2w)
            for (int32_t t0 = 0; t0 < deg; t0 += 64) {
                int32_t c1 = -1;
                const int32_t t = t0 + lane;
                if (t < deg) {
                    const int32_t ul = a.adjncy[beg + t] - v0;
                    const int32_t w = a.adjwgt ? a.adjwgt[beg + t] : 1;
                    const uint8_t su = state[ul];
                    if (!(su & kLocked)) {
                        const bool same = (su & kSide) == from;
                        const int32_t w2 = wadd(w, w);
                        atomicAdd(&gains[ul], same ? w2 : wsub(0, w2));
                        atomicAdd(&exts[ul], same ? 1 : -1);
                        c1 = ul;
                    }
                }
                wave_fence_sync();
                bool q = false;
                uint8_t su = 0;
                if (c1 >= 0) {
                    su = state[c1];
                    q = (su & kQueued) || exts[c1] > 0;
                    if (q) leaves[c1] = f2_key(gains[c1], c1);
                }
                wave_fence_sync();
                if (q) state[c1] = su | kQueued;
                f2_tree_update(nlvl, s_len, s_off, leaves, tree, c0, q ? c1 : -1);
                c0 = -1;
                pending = false;
            }
        } else if (feas) {
            // the unlocked neighbours' gains from their own rows (fm.rs:99-118)
            for (int32_t t0 = 0; t0 < deg; t0 += 64) {
                int32_t c1 = -1;
                const int32_t t = t0 + lane;
                if (t < deg) {
                    const int32_t ul = a.adjncy[beg + t] - v0;
                    const uint8_t su = state[ul];
                    if (!(su & kLocked)) {
                        const int32_t pu = su & kSide;
                        const int32_t ub = a.xadj[v0 + ul], ue = a.xadj[v0 + ul + 1];
                        int32_t gu = 0;
                        bool ext = false;
                        int32_t j = ub;
                        for (; j + 4 <= ue; j += 4) {
                            int32_t x[4], w[4];
#pragma unroll
                            for (int q = 0; q < 4; ++q) x[q] = a.adjncy[j + q];
#pragma unroll
                            for (int q = 0; q < 4; ++q) w[q] = a.adjwgt ? a.adjwgt[j + q] : 1;
#pragma unroll
                            for (int q = 0; q < 4; ++q) {
                                if ((state[x[q] - v0] & kSide) == pu) {
                                    gu = wsub(gu, w[q]);
                                } else {
                                    gu = wadd(gu, w[q]);
                                    ext = true;
                                }
                            }
                        }
                        for (; j < ue; ++j) {
                            const int32_t w = a.adjwgt ? a.adjwgt[j] : 1;
                            if ((state[a.adjncy[j] - v0] & kSide) == pu) {
                                gu = wsub(gu, w);
                            } else {
                                gu = wadd(gu, w);
                                ext = true;
                            }
                        }
                        if ((su & kQueued) || ext) {
                            leaves[ul] = f2_key(gu, ul);
                            c1 = ul;
                        }
                    }
                }
                wave_fence_sync();
                if (c1 >= 0) state[c1] |= kQueued;
                f2_tree_update(nlvl, s_len, s_off, leaves, tree, c0, c1);
                c0 = -1;
                pending = false;
            }
        }
        if (pending) f2_tree_update(nlvl, s_len, s_off, leaves, tree, c0, -1);
    }
    wave_fence_sync();
    // the moved vertices' sides back to the part array
    for (int32_t k = lane; k < nm; k += 64) {
        const int32_t i = a.moves[v0 + k];
        a.part[v0 + i] = state[i] & kSide;
    }
    for (int k = lane; k < 2 * ncon; k += 64) pwg[k] = s_pw[k];
    if (lane == 0) {
        a.best_len[gd.g] = best;
        a.best_cut[gd.g] = best_cut;
        a.nmoves[gd.g] = nm;
    }
}

// The undo (fm.rs:121-131) and the pass's verdict for graph blockIdx.x.
__global__ __launch_bounds__(kF2Threads) void k_fm2_undo(Fm2Args a) {
    __shared__ int32_t s_d[2 * kF2MaxCon];
    const Fm2Graph gd = a.graphs[blockIdx.x];
    if (a.done[gd.g]) return;
    const int tid = threadIdx.x, ncon = a.ncon;
    const int32_t best = a.best_len[gd.g], nm = a.nmoves[gd.g];
    for (int i = tid; i < 2 * ncon; i += kF2Threads) s_d[i] = 0;
    __syncthreads();
    // the moves past the best prefix are distinct vertices: each goes back
    for (int32_t k = best + tid; k < nm; k += kF2Threads) {
        const int32_t v = gd.v0 + a.moves[gd.v0 + k];
        const int32_t cur = a.part[v], from = 1 - cur;
        a.part[v] = from;
        for (int c = 0; c < ncon; ++c) {
            const int32_t vw = a.vwgt ? a.vwgt[(int64_t)v * ncon + c] : 1;
            atomicAdd(&s_d[from * ncon + c], vw);
            atomicSub(&s_d[cur * ncon + c], vw);
        }
    }
    __syncthreads();
    int32_t* pwg = a.pw + a.gkc[gd.g];
    for (int i = tid; i < 2 * ncon; i += kF2Threads) pwg[i] = wadd(pwg[i], s_d[i]);
    if (tid == 0) {
        const bool improved = best > 0 && a.best_cut[gd.g] < a.cut1[gd.g];   // fm.rs:122
        if (!improved) a.done[gd.g] = 1;                                      // fm.rs:23-25
    }
}

// *flag = 0 unless every entry (v, u, w) of the graph has as many copies
// in v's row as (v, w) has in u's row (the rows are then pair-symmetric).
// Vertex pairs whose rows together exceed kSymMaxScan entries are not
// examined and clear the flag as well (the exact general path is used).
constexpr int32_t kSymMaxScan = 4096;
__global__ __launch_bounds__(kBlock) void k_pair_symmetry(int32_t n, const int32_t* __restrict__ xadj,
                                                          const int32_t* __restrict__ adjncy,
                                                          const int32_t* __restrict__ adjwgt,
                                                          int32_t* __restrict__ flag) {
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock) {
        const int32_t vb = xadj[v], ve = xadj[v + 1];
        bool ok = true;
        for (int32_t j = vb; j < ve && ok; ++j) {
            const int32_t u = adjncy[j], w = adjwgt ? adjwgt[j] : 1;
            const int32_t ub = xadj[u], ue = xadj[u + 1];
            if ((ve - vb) + (ue - ub) > kSymMaxScan) {
                ok = false;
                break;
            }
            int32_t c1 = 0, c2 = 0;
            for (int32_t k = vb; k < ve; ++k) c1 += (adjncy[k] == u && (adjwgt ? adjwgt[k] : 1) == w) ? 1 : 0;
            for (int32_t k = ub; k < ue; ++k) c2 += (adjncy[k] == (int32_t)v && (adjwgt ? adjwgt[k] : 1) == w) ? 1 : 0;
            ok = c1 == c2;
        }
        if (!ok) *flag = 0;
    }
}

inline int grid_of(int64_t items, int per_block) {
    int64_t b = (items + per_block - 1) / per_block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 16384));
}

}  // namespace

void fm2_symmetry(const DevGraph& g, int32_t* flag, hipStream_t s) {
    const int32_t one = 1;
    RK_HIP(hipMemcpyAsync(flag, &one, sizeof(int32_t), hipMemcpyHostToDevice, s));
    RK_HIP(hipStreamSynchronize(s));
    if (g.n > 0)
        hipLaunchKernelGGL(k_pair_symmetry, dim3(grid_of(g.n, kBlock)), dim3(kBlock), 0, s, g.n, g.xadj, g.adjncy,
                           g.adjwgt, flag);
    RK_HIP(hipGetLastError());
}

int64_t fm2_tree_nodes(int32_t n) {
    int64_t t = 0, len = n;
    while (len > 1) {
        len = (len + kF2Fan - 1) / kF2Fan;
        t += len;
    }
    return t;
}

int32_t fm2_tree_levels(int32_t n) {
    int l = 1;
    for (int64_t len = n; len > 1; len = (len + kF2Fan - 1) / kF2Fan) ++l;
    return l;
}

int64_t fm2_level_nodes(int32_t n, int l) {
    int64_t len = n;
    for (int j = 0; j < l && len > 1; ++j) len = (len + kF2Fan - 1) / kF2Fan;
    return l < fm2_tree_levels(n) ? len : 0;
}

void fm2_refine_batch(const DevGraph& g, const Fm2Level& lv, Fm2Bufs& f, const int32_t* gkc, const long long* fmax,
                      int32_t* pw, int32_t* part, int32_t* gdone, int niter, unsigned long long* cutacc,
                      hipStream_t s) {
    if (lv.ngraphs <= 0) return;
    if (g.ncon > kF2MaxCon) throw Error(-2, "2-way FM refinement supports at most 16 constraints");
    if (lv.nlvl > kFm2MaxLvl) throw Error(-4, "2-way FM queue tree deeper than its table");
    Fm2Args a{};
    a.ncon = g.ncon;
    a.xadj = g.xadj;
    a.adjncy = g.adjncy;
    a.adjwgt = g.adjwgt;
    a.vwgt = g.vwgt;
    a.part = part;
    a.pw = pw;
    a.fmax = fmax;
    a.gkc = gkc;
    a.gseg = lv.gseg;
    a.gdesc_of = lv.gdesc_of;
    a.graphs = lv.graphs;
    a.ngraphs = lv.ngraphs;
    a.leaf = f.leaf;
    a.tree = f.tree;
    a.state = f.state;
    a.gain = f.gain;
    a.ext = f.ext;
    a.sym = f.sym;
    a.moves = f.moves;
    a.cut0 = f.cut0;
    a.cut1 = f.cut1;
    a.best_len = f.best_len;
    a.best_cut = f.best_cut;
    a.nmoves = f.nmoves;
    a.done = gdone;
    a.err = f.err;
    static const hipError_t lds_ok =
        hipFuncSetAttribute((const void*)k_fm2_pass<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kF2Lds);
    static const hipError_t lds_ok1 =
        hipFuncSetAttribute((const void*)k_fm2_pass<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kF2Lds);
    RK_HIP(lds_ok);
    RK_HIP(lds_ok1);
    for (int it = 0; it < niter; ++it) {                                  // fm.rs:22-26
        hipLaunchKernelGGL(k_fm2_init, dim3(grid_of(g.n, kBlock / kF2G)), dim3(kBlock), 0, s, a, g.n);
        for (int l = 1; l < lv.nlvl; ++l)
            hipLaunchKernelGGL(k_fm2_tree, dim3(grid_of(lv.level_nodes[l], kBlock)), dim3(kBlock), 0, s, a, l,
                               lv.nodes + (int64_t)l * (lv.ngraphs + 1));
        edge_cut_batch(g, part, lv.gseg, lv.gtotal, cutacc, f.cut0, s);   // fm.rs:60
        hipLaunchKernelGGL(k_fm2_pass<2>, dim3(lv.ngraphs), dim3(kF2Threads), kF2Lds, s, a);
        if (!fm2_fits_lds(lv.max_n)) {
            hipLaunchKernelGGL(k_fm2_pass<1>, dim3(lv.ngraphs), dim3(kF2Threads), kF2Lds, s, a);
            hipLaunchKernelGGL(k_fm2_pass<0>, dim3(lv.ngraphs), dim3(kF2Threads), 0, s, a);
        }
        edge_cut_batch(g, part, lv.gseg, lv.gtotal, cutacc, f.cut1, s);   // fm.rs:122
        hipLaunchKernelGGL(k_fm2_undo, dim3(lv.ngraphs), dim3(kF2Threads), 0, s, a);
    }
    RK_HIP(hipGetLastError());
}

}  // namespace rk
