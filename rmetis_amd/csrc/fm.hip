// fm.hip -- k-way FM refinement (refine/fm_kway.rs:12-158) on the GPU.
//
// A pass of the reference builds a priority queue of the boundary vertices
// with a positive best gain (fm_kway.rs:47-85), then pops the best vertex,
// moves it to its best part when the move keeps that part within its limit
// (105-109), locks it and recomputes the best move of every unlocked
// neighbour (126-143), until the queue is empty.  Every kept move has a
// positive gain, so the cut falls with each one and the rollback to the best
// prefix (146-155) keeps them all.
//
// The pops are inherently ordered: each decision reads the part weights and
// the neighbour gains the earlier pops left.  Here:
//  * the initial queue is built in parallel: the gain-table kernel lists
//    every boundary vertex's parts with a nonzero edge-weight sum in (gain
//    descending, part ascending) order, whose head is the reference's best
//    move (62-70); the queue's keys (gain, then the smaller vertex first --
//    the order among equal gains that this build fixes) become the leaves of
//    a 64-ary max tree, built one level per launch;
//  * one workgroup then runs the pops: the root names the vertex, one thread
//    decides the move, the threads of the workgroup recompute the moved
//    vertex's neighbours (their distinct parts in registers, or a dense part
//    table for vertices with many), and the waves rebuild the tree nodes
//    above the changed leaves, one level per barrier.
#include "common.hpp"
#include "kernels.hpp"

namespace rk {

namespace {

constexpr int kFmThreads = 256;   // the workgroup of the pops
constexpr int kFmWaves = kFmThreads / 64;
constexpr int kFmSlots = 8;       // distinct neighbour parts held in registers by one recompute
constexpr int kFmLdsPw = 4096;    // part-weight entries held in LDS
constexpr int kFmLdsTab = 4096;   // parts of the dense table in LDS (larger k: global)
constexpr int kFmChg = 2048;      // changed leaves tracked per pop (more: the tree is rebuilt)
constexpr int kFmMaxLvl = 8;

struct FmArgs {
    int32_t n, ncon, k;
    const int32_t* xadj;
    const int32_t* adjncy;
    const int32_t* adjwgt;
    const int32_t* vwgt;
    int32_t* part;
    int32_t* pw;
    const long long* fmax;
    uint64_t* tree;
    int32_t nlvl;
    int64_t off[kFmMaxLvl];
    int64_t len[kFmMaxLvl];
    int32_t* tp;
    uint8_t* locked;
    int32_t* done;
    int32_t* err;
    int32_t* ovf;
    const int4* vrec;
    const int32_t* pool;
    const int32_t* gains;
};

__device__ inline uint64_t fm_key(int32_t gain, int32_t v) {
    return ((uint64_t)(uint32_t)gain << 32) | (uint32_t)(0x7FFFFFFF - v);
}

// Leaves of the queue: the boundary vertices whose best move (the head of
// their candidate list) has a positive gain.  Also clears the locks.
__global__ __launch_bounds__(kBlock) void k_fm_leaves(FmArgs a) {
    if (*a.done) return;
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < a.n; v += (int64_t)gridDim.x * kBlock) {
        a.locked[v] = 0;
        const int4 rec = a.vrec[v];
        uint64_t key = 0;
        if (rec.x >= 0 && rec.z > 0) {
            const int32_t j = a.xadj[v];
            const int32_t g = a.gains[j];
            if (g > 0) {
                key = fm_key(g, (int32_t)v);
                a.tp[v] = a.pool[j];
            }
        }
        a.tree[v] = key;
    }
}

__device__ inline uint64_t wave_max_u64(uint64_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t y = __shfl_xor(x, o, 64);
        x = y > x ? y : x;
    }
    return x;
}

// Level l of the tree from level l - 1: one wave per node.
__global__ __launch_bounds__(kBlock) void k_fm_tree_level(FmArgs a, int l) {
    if (*a.done) return;
    const int lane = __lane_id();
    const int64_t nodes = a.len[l];
    for (int64_t i = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / 64; i < nodes;
         i += (int64_t)gridDim.x * (kBlock / 64)) {
        const int64_t c = i * 64 + lane;
        uint64_t x = c < a.len[l - 1] ? a.tree[a.off[l - 1] + c] : 0;
        x = wave_max_u64(x);
        if (lane == 0) a.tree[a.off[l] + i] = x;
    }
}

// Best move of u against the current parts (fm_kway.rs:47-72): the part
// p != part[u] with a nonzero edge-weight sum of largest gain sum_p - sum_own,
// the lower part on ties, when that gain is positive.  Up to kFmSlots
// distinct other parts in registers; *more is set when u has more.
__device__ inline bool fm_best_regs(const FmArgs& a, int32_t u, int32_t* bp, int32_t* bg, bool* more) {
    int32_t sp[kFmSlots], sw[kFmSlots];
#pragma unroll
    for (int t = 0; t < kFmSlots; ++t) {
        sp[t] = -1;
        sw[t] = 0;
    }
    const int32_t pu = a.part[u];
    int32_t own = 0;
    for (int32_t j = a.xadj[u]; j < a.xadj[u + 1]; ++j) {
        const int32_t q = a.part[a.adjncy[j]];
        const int32_t w = a.adjwgt ? a.adjwgt[j] : 1;
        if (q == pu) {
            own = wadd(own, w);
            continue;
        }
        bool placed = false;
#pragma unroll
        for (int t = 0; t < kFmSlots; ++t) {
            if (!placed && sp[t] == q) {
                sw[t] = wadd(sw[t], w);
                placed = true;
            }
        }
#pragma unroll
        for (int t = 0; t < kFmSlots; ++t) {
            if (!placed && sp[t] < 0) {
                sp[t] = q;
                sw[t] = w;
                placed = true;
            }
        }
        if (!placed) {
            *more = true;
            return false;
        }
    }
    int32_t best_p = -1, best_g = 0;
#pragma unroll
    for (int t = 0; t < kFmSlots; ++t) {
        if (sp[t] < 0 || sw[t] == 0) continue;
        const int32_t g = wsub(sw[t], own);
        if (g > best_g || (g == best_g && best_p >= 0 && sp[t] < best_p)) {
            best_g = g;
            best_p = sp[t];
        }
    }
    if (best_p < 0) return false;
    *bp = best_p;
    *bg = best_g;
    return true;
}

// The same by one wave over a dense table of k part sums (tab: LDS or global).
__device__ inline bool fm_best_wave(const FmArgs& a, int32_t u, int32_t* tab, int32_t* bp, int32_t* bg) {
    const int lane = __lane_id();
    for (int p = lane; p < a.k; p += 64) tab[p] = 0;
    wave_sync();
    const int32_t beg = a.xadj[u], end = a.xadj[u + 1];
    for (int32_t j = beg + lane; j < end; j += 64) atomicAdd(&tab[a.part[a.adjncy[j]]], a.adjwgt ? a.adjwgt[j] : 1);
    wave_sync();
    const int32_t pu = a.part[u];
    const int32_t own = tab[pu];
    int32_t best_p = 0x7FFFFFFF, best_g = 0;
    for (int p = lane; p < a.k; p += 64) {
        const int32_t val = tab[p];
        if (p == pu || val == 0) continue;
        const int32_t g = wsub(val, own);
        if (g > best_g || (g == best_g && best_g > 0 && p < best_p)) {
            best_g = g;
            best_p = p;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int32_t og = __shfl_xor(best_g, o, 64), op = __shfl_xor(best_p, o, 64);
        if (og > best_g || (og == best_g && op < best_p)) {
            best_g = og;
            best_p = op;
        }
    }
    wave_sync();
    if (best_g <= 0) return false;
    *bp = best_p;
    *bg = best_g;
    return true;
}

// The pops of one pass (fm_kway.rs:93-144), one workgroup.
__global__ __launch_bounds__(kFmThreads) void k_fm_pass(FmArgs a) {
    __shared__ int32_t s_pw[kFmLdsPw];
    __shared__ int32_t s_tab[kFmLdsTab];
    __shared__ int32_t s_chg[kFmChg];
    __shared__ int32_t s_ovf[kFmThreads];
    __shared__ int32_t s_cmd[4];
    __shared__ int32_t s_nchg, s_novf;
    if (*a.done) return;
    const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
    const int kc = a.k * a.ncon;
    const bool lds_pw = kc <= kFmLdsPw;
    int32_t* pw = lds_pw ? s_pw : a.pw;
    if (lds_pw)
        for (int i = tid; i < kc; i += kFmThreads) s_pw[i] = a.pw[i];
    int32_t* tab = a.k <= kFmLdsTab ? s_tab : a.ovf;
    const int64_t root = a.off[a.nlvl - 1];
    // every pop removes a queued vertex, and a vertex is queued again only
    // by a move of one of its neighbours: at most n + m pops
    int64_t budget = (int64_t)a.n + a.xadj[a.n] + 1;
    int32_t moves = 0;
    __syncthreads();
    for (;;) {
        if (tid == 0) {
            const uint64_t key = a.tree[root];
            int32_t v = -1, to = 0, from = 0, feas = 0;
            if (key != 0) {
                v = 0x7FFFFFFF - (int32_t)(uint32_t)key;
                to = a.tp[v];
                from = a.part[v];
                feas = 1;
                for (int c = 0; c < a.ncon; ++c) {                        // fm_kway.rs:105-109
                    const int32_t vw = a.vwgt ? a.vwgt[(int64_t)v * a.ncon + c] : 1;
                    const int32_t nw = wadd(pw[to * a.ncon + c], vw);
                    if ((long long)nw > a.fmax[to * a.ncon + c]) feas = 0;
                }
                a.tree[v] = 0;                                            // popped
                if (feas) {                                               // fm_kway.rs:112-118
                    a.part[v] = to;
                    for (int c = 0; c < a.ncon; ++c) {
                        const int32_t vw = a.vwgt ? a.vwgt[(int64_t)v * a.ncon + c] : 1;
                        pw[from * a.ncon + c] = wsub(pw[from * a.ncon + c], vw);
                        pw[to * a.ncon + c] = wadd(pw[to * a.ncon + c], vw);
                    }
                    a.locked[v] = 1;
                }
            }
            s_cmd[0] = v;
            s_cmd[1] = feas;
            s_chg[0] = v;
            s_nchg = 1;
            s_novf = 0;
        }
        __syncthreads();
        const int32_t v = s_cmd[0];
        if (v < 0) break;
        if (--budget < 0) {
            if (tid == 0) *a.err = 1;
            break;
        }
        if (s_cmd[1]) {
            ++moves;
            const int32_t beg = a.xadj[v], deg = a.xadj[v + 1] - beg;
            for (int32_t t0 = 0; t0 < deg; t0 += kFmThreads) {            // fm_kway.rs:127-143
                const int32_t t = t0 + tid;
                if (t < deg) {
                    const int32_t u = a.adjncy[beg + t];
                    if (!a.locked[u]) {
                        int32_t bp = 0, bg = 0;
                        bool more = false;
                        const bool some = fm_best_regs(a, u, &bp, &bg, &more);
                        if (more) {
                            s_ovf[atomicAdd(&s_novf, 1)] = u;
                        } else {
                            a.tree[u] = some ? fm_key(bg, u) : 0;
                            if (some) a.tp[u] = bp;
                            const int c = atomicAdd(&s_nchg, 1);
                            if (c < kFmChg) s_chg[c] = u;
                        }
                    }
                }
                __syncthreads();
                // neighbours with more distinct parts than registers hold:
                // wave 0 recomputes them one by one over the dense table
                if (wave == 0) {
                    const int no = s_novf;
                    for (int i = 0; i < no; ++i) {
                        const int32_t u = s_ovf[i];
                        int32_t bp = 0, bg = 0;
                        const bool some = fm_best_wave(a, u, tab, &bp, &bg);
                        if (lane == 0) {
                            a.tree[u] = some ? fm_key(bg, u) : 0;
                            if (some) a.tp[u] = bp;
                            const int c = atomicAdd(&s_nchg, 1);
                            if (c < kFmChg) s_chg[c] = u;
                        }
                    }
                }
                __syncthreads();
                if (tid == 0) s_novf = 0;
                __syncthreads();
            }
        }
        // the tree above the changed leaves, one level per barrier; past
        // kFmChg changes every node is rebuilt.  Changed leaves are mostly
        // neighbours that share parents: each distinct node of a level is
        // rebuilt once, by the wave of its first entry (the lanes compare
        // the earlier entries' nodes, 64 at a time).
        const int nchg = s_nchg;
        const bool all = nchg > kFmChg;
        for (int l = 1; l < a.nlvl; ++l) {
            const int64_t cnt = all ? a.len[l] : nchg;
            for (int64_t i = wave; i < cnt; i += kFmWaves) {
                if (!all) {
                    const int32_t nd = s_chg[i] >> (6 * l);
                    bool dup = false;
                    for (int64_t j0 = 0; j0 < i && !dup; j0 += 64) {
                        const int64_t j = j0 + lane;
                        dup = __ballot(j < i && (s_chg[j] >> (6 * l)) == nd) != 0;
                    }
                    if (dup) continue;
                }
                const int64_t node = all ? i : ((int64_t)s_chg[i] >> (6 * l));
                const int64_t c = node * 64 + lane;
                uint64_t x = c < a.len[l - 1] ? a.tree[a.off[l - 1] + c] : 0;
                x = wave_max_u64(x);
                if (lane == 0) a.tree[a.off[l] + node] = x;
            }
            __syncthreads();
        }
    }
    if (lds_pw) {
        __syncthreads();
        for (int i = tid; i < kc; i += kFmThreads) a.pw[i] = s_pw[i];
    }
    if (tid == 0 && moves == 0) *a.done = 1;                             // fm_kway.rs:24: a pass without a move ends
}

__global__ void k_fm_fill(int32_t* pw0, long long* omin, int64_t kc) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < kc; i += (int64_t)gridDim.x * blockDim.x) {
        pw0[i] = 0;
        omin[i] = LLONG_MIN;
    }
}

}  // namespace

void fm_fill_tables(int32_t* pw0, long long* omin, int64_t kc, hipStream_t s) {
    const int grid = (int)std::min<int64_t>((kc + kBlock - 1) / kBlock + 1, 1024);
    hipLaunchKernelGGL(k_fm_fill, dim3(grid), dim3(kBlock), 0, s, pw0, omin, kc);
    RK_HIP(hipGetLastError());
}

void fm_refine_level(const DevGraph& g, int32_t k, const long long* d_fmax, int niter, int32_t* part, int nheavy,
                     const int32_t* heavy_list, RefineBufs& b, FmBufs& f, hipStream_t s, bool cleared,
                     bool pw_valid) {
    if (g.n <= 0) return;
    if (!pw_valid) part_weights(g, part, k, b.pw, s, cleared);            // fm_kway.rs:21
    if (!cleared) RK_HIP(hipMemsetAsync(b.done, 0, sizeof(int32_t), s));
    FmArgs a{};
    a.n = g.n;
    a.ncon = g.ncon;
    a.k = k;
    a.xadj = g.xadj;
    a.adjncy = g.adjncy;
    a.adjwgt = g.adjwgt;
    a.vwgt = g.vwgt;
    a.part = part;
    a.pw = b.pw;
    a.fmax = d_fmax;
    a.tree = f.tree;
    int64_t len = g.n, off = 0;
    int nl = 0;
    for (;;) {
        a.off[nl] = off;
        a.len[nl] = len;
        ++nl;
        if (len <= 1) break;
        off += len;
        len = (len + 63) / 64;
    }
    if (nl > kFmMaxLvl) throw Error(-4, "FM queue tree deeper than its table");
    a.nlvl = nl;
    a.tp = f.tp;
    a.locked = f.locked;
    a.done = b.done;
    a.err = f.err;
    a.ovf = f.ovf;
    a.vrec = f.vrec;
    a.pool = b.pool;
    a.gains = f.gains;
    const int lgrid = (int)std::min<int64_t>((g.n + kBlock - 1) / kBlock, 4096);
    for (int it = 0; it < niter; ++it) {                                  // fm_kway.rs:23-27
        fm_candidates(g, k, nheavy, heavy_list, b, part, f.pw0, f.omin, b.done, f.vrec, f.g0, f.gains, s);
        hipLaunchKernelGGL(k_fm_leaves, dim3(lgrid), dim3(kBlock), 0, s, a);
        for (int l = 1; l < nl; ++l) {
            const int grid = (int)std::min<int64_t>((a.len[l] * 64 + kBlock - 1) / kBlock, 4096);
            hipLaunchKernelGGL(k_fm_tree_level, dim3(grid), dim3(kBlock), 0, s, a, l);
        }
        hipLaunchKernelGGL(k_fm_pass, dim3(1), dim3(kFmThreads), 0, s, a);
    }
    RK_HIP(hipGetLastError());
}

}  // namespace rk
