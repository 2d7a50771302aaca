// coarsen.hip -- matching and coarse-graph contraction on the GPU.
//
// Matching (coarsen/shem.rs:13-56, coarsen/rm.rs:13-43).  The reference's
// matchings are sequential greedy scans.  On the GPU every unmatched vertex
// proposes to its best unmatched neighbour under a strict total order on
// edges -- (edge weight, seeded edge hash) for heavy-edge matching, the hash
// alone for random matching -- and mutual proposals become pairs (a
// locally-dominant handshake).  Rounds repeat until nothing changes or the
// round limit; leftovers are self-matched.  The result is symmetric,
// deterministic for a given seed and independent of scheduling.
//
// Contraction (graph/mod.rs:284-384), bit-exact for a given symmetric matching:
//  * coarse ids ascend with the pair's smaller vertex (288-297): exclusive
//    scan of leader flags match[v] >= v;
//  * coarse vertex weights are summed, and stay absent when the fine graph
//    has none (300-314);
//  * coarse adjacency of {a < b}: first occurrence over adj(a) then adj(b)
//    in CSR order, internal edges dropped, weights of merged entries summed
//    (323-341).  Small coarse vertices use a per-group LDS hash that records
//    each coarse neighbour's first position and weight sum; a second pass
//    emits first occurrences with a ballot prefix.  Large ones (degree sum
//    > 8G) use a block per vertex and a hash table in global memory.
#include "common.hpp"
#include "kernels.hpp"

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

// Strict total order on edges, symmetric in (v, u).  Heavy-edge matching:
// (smaller endpoint degree first, heavier edge, seeded hash) -- an edge is
// ranked by when the reference's degree-ordered scan (shem.rs:23) would
// first reach it and by the weight its visiting endpoint maximises
// (shem.rs:36-44).  Random matching: the seeded hash alone.
struct EdgeKey {
    uint64_t hi;
    uint32_t lo;
    __device__ bool gt(const EdgeKey& o) const { return hi > o.hi || (hi == o.hi && lo > o.lo); }
    __device__ bool eq(const EdgeKey& o) const { return hi == o.hi && lo == o.lo; }
};

__device__ inline EdgeKey edge_key(int32_t v, int32_t u, int32_t w, int32_t dv, int32_t du, uint64_t seed,
                                   int shem) {
    const uint32_t lo = (uint32_t)min(v, u), hi = (uint32_t)max(v, u);
    const uint32_t h = mix32(lo ^ mix32(hi ^ (uint32_t)seed) ^ (uint32_t)(seed >> 32));
    if (!shem) return EdgeKey{0ull, h};
    const uint32_t dpri = 0xFFFFFFFFu - (uint32_t)min(dv, du);
    return EdgeKey{((uint64_t)dpri << 32) | (uint64_t)((uint32_t)w ^ 0x80000000u), h};
}

__global__ __launch_bounds__(kBlock) void k_match_init(int32_t n, const int32_t* __restrict__ xadj,
                                                       int32_t* __restrict__ match) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v < n) match[v] = (xadj[v + 1] == xadj[v]) ? (int32_t)v : -1;
}

// Round r runs only if round r-1 matched or retired something (flags[r-1]).
template <int G>
__global__ __launch_bounds__(kBlock) void k_match_propose(int32_t n, const int32_t* __restrict__ xadj,
                                                          const int32_t* __restrict__ adjncy,
                                                          const int32_t* __restrict__ adjwgt,
                                                          const int32_t* __restrict__ match,
                                                          int32_t* __restrict__ cand, uint64_t seed, int shem,
                                                          const int32_t* __restrict__ prev_changed) {
    if (prev_changed && *prev_changed == 0) return;
    constexpr int NG = kBlock / G;
    const int gl = group_lane<G>();
    const int64_t v = (int64_t)blockIdx.x * NG + threadIdx.x / G;
    if (v >= n) return;
    if (match[v] != -1) return;   // group-uniform
    const int32_t beg = xadj[v], end = xadj[v + 1];
    const int32_t dv = end - beg;
    EdgeKey best{0ull, 0u};
    int32_t bu = -1;
    for (int32_t j = beg + gl; j < end; j += G) {
        const int32_t u = adjncy[j];
        if (match[u] != -1) continue;
        const int32_t du = xadj[u + 1] - xadj[u];
        const EdgeKey key = edge_key((int32_t)v, u, adjwgt ? adjwgt[j] : 1, dv, du, seed, shem);
        if (bu < 0 || key.gt(best) || (key.eq(best) && u < bu)) {
            best = key;
            bu = u;
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        const EdgeKey ob{__shfl_xor(best.hi, o, G), __shfl_xor(best.lo, o, G)};
        const int32_t ou = __shfl_xor(bu, o, G);
        if (ou >= 0 && (bu < 0 || ob.gt(best) || (ob.eq(best) && ou < bu))) {
            best = ob;
            bu = ou;
        }
    }
    if (gl == 0) cand[v] = bu >= 0 ? bu : -2;
}

__global__ __launch_bounds__(kBlock) void k_match_accept(int32_t n, const int32_t* __restrict__ cand,
                                                         int32_t* __restrict__ match, int32_t* __restrict__ changed,
                                                         const int32_t* __restrict__ prev_changed) {
    if (prev_changed && *prev_changed == 0) return;
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    bool ch = false;
    if (v < n && match[v] == -1) {
        const int32_t u = cand[v];
        if (u == -2) {
            match[v] = (int32_t)v;
            ch = true;
        } else if (u >= 0 && cand[u] == (int32_t)v) {
            match[v] = u;
            ch = true;
        }
    }
    if (__ballot(ch) && __lane_id() == 0) atomicOr(changed, 1);
}

__global__ __launch_bounds__(kBlock) void k_match_finish(int32_t n, int32_t* __restrict__ match) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v < n && match[v] == -1) match[v] = (int32_t)v;
}

__global__ __launch_bounds__(kBlock) void k_leader(int32_t n, const int32_t* __restrict__ match,
                                                   int32_t* __restrict__ leader) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v < n) leader[v] = match[v] >= (int32_t)v ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_cmap(int32_t n, int32_t ncon, const int32_t* __restrict__ xadj,
                                                 const int32_t* __restrict__ vwgt, const int32_t* __restrict__ vsize,
                                                 const int32_t* __restrict__ match, const int32_t* __restrict__ cid,
                                                 int32_t* __restrict__ cmap, int32_t* __restrict__ members,
                                                 int32_t* __restrict__ degsum, int32_t* __restrict__ cvwgt,
                                                 int32_t* __restrict__ cvsize) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v >= n) return;
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
    }
}

// Position t of the merged adjacency sequence adj(a) ++ adj(b) -> CSR index.
__device__ inline int32_t seq_edge(const int32_t* xadj, int32_t a, int32_t b, int32_t dega, int32_t t) {
    return t < dega ? xadj[a] + t : xadj[b] + (t - dega);
}

__device__ inline uint32_t hslot(int32_t key, uint32_t mask) { return (mix32((uint32_t)key) & mask); }

template <int G>
__global__ __launch_bounds__(kBlock) void k_contract(int32_t cn, const int32_t* __restrict__ xadj,
                                                     const int32_t* __restrict__ adjncy,
                                                     const int32_t* __restrict__ adjwgt,
                                                     const int32_t* __restrict__ cmap,
                                                     const int32_t* __restrict__ members,
                                                     const int32_t* __restrict__ degsum,
                                                     const int32_t* __restrict__ tmp_off,
                                                     int32_t* __restrict__ tmp_adj, int32_t* __restrict__ tmp_w,
                                                     int32_t* __restrict__ cdeg, int32_t* __restrict__ big_list,
                                                     int32_t* __restrict__ counters) {
    constexpr int NG = kBlock / G;
    constexpr int S = 8 * G;
    __shared__ int32_t s_key[NG * S];
    __shared__ int32_t s_first[NG * S];
    __shared__ int32_t s_sum[NG * S];
    const int grp = threadIdx.x / G;
    const int gl = group_lane<G>();
    int32_t* hk = s_key + grp * S;
    int32_t* hf = s_first + grp * S;
    int32_t* hs = s_sum + grp * S;
    for (int s = gl; s < S; s += G) {
        hk[s] = -1;
        hf[s] = INT32_MAX;
        hs[s] = 0;
    }
    wave_sync();
    bool not_one = false;
    for (int64_t cv = (int64_t)blockIdx.x * NG + grp; cv < cn; cv += (int64_t)gridDim.x * NG) {
        const int32_t ds = degsum[cv];
        if (ds > S) {
            if (gl == 0) big_list[atomicAdd(&counters[2], 1)] = (int32_t)cv;
            continue;
        }
        const int32_t a = members[2 * cv], b = members[2 * cv + 1];
        const int32_t dega = xadj[a + 1] - xadj[a];
        for (int32_t t = gl; t < ds; t += G) {
            const int32_t j = seq_edge(xadj, a, b, dega, t);
            const int32_t cu = cmap[adjncy[j]];
            if (cu == (int32_t)cv) continue;
            const int32_t w = adjwgt ? adjwgt[j] : 1;
            uint32_t s = hslot(cu, S - 1);
            for (;;) {
                const int32_t prev = atomicCAS(&hk[s], -1, cu);
                if (prev == -1 || prev == cu) {
                    atomicMin(&hf[s], t);
                    atomicAdd(&hs[s], w);
                    break;
                }
                s = (s + 1) & (S - 1);
            }
        }
        wave_sync();
        const int32_t off = tmp_off[cv];
        int32_t carry = 0;
        for (int32_t t0 = 0; t0 < ds; t0 += G) {
            const int32_t t = t0 + gl;
            bool first = false;
            int32_t cu = -1, sum = 0;
            if (t < ds) {
                const int32_t j = seq_edge(xadj, a, b, dega, t);
                cu = cmap[adjncy[j]];
                if (cu != (int32_t)cv) {
                    uint32_t s = hslot(cu, S - 1);
                    while (hk[s] != cu) s = (s + 1) & (S - 1);
                    first = hf[s] == t;
                    sum = hs[s];
                }
            }
            const uint64_t bm = group_ballot<G>(first);
            if (first) {
                const int pos = carry + __popcll(bm & ((1ull << gl) - 1ull));
                tmp_adj[off + pos] = cu;
                tmp_w[off + pos] = sum;
                if (sum != 1) not_one = true;
            }
            carry += __popcll(bm);
        }
        if (gl == 0) cdeg[cv] = carry;
        wave_sync();
        for (int s = gl; s < S; s += G) {
            hk[s] = -1;
            hf[s] = INT32_MAX;
            hs[s] = 0;
        }
        wave_sync();
    }
    if (__ballot(not_one) && __lane_id() == 0) atomicOr(&counters[4], 1);
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

template <int G>
__global__ __launch_bounds__(kBlock) void k_contract_copy(int32_t cn, const int32_t* __restrict__ tmp_off,
                                                          const int32_t* __restrict__ tmp_adj,
                                                          const int32_t* __restrict__ tmp_w,
                                                          const int32_t* __restrict__ cxadj,
                                                          int32_t* __restrict__ cadj, int32_t* __restrict__ cw) {
    constexpr int NG = kBlock / G;
    const int gl = group_lane<G>();
    for (int64_t cv = (int64_t)blockIdx.x * NG + threadIdx.x / G; cv < cn; cv += (int64_t)gridDim.x * NG) {
        const int32_t src = tmp_off[cv];
        const int32_t dst = cxadj[cv];
        const int32_t cnt = cxadj[cv + 1] - dst;
        for (int32_t i = gl; i < cnt; i += G) {
            cadj[dst + i] = tmp_adj[src + i];
            cw[dst + i] = tmp_w[src + i];
        }
    }
}

inline int grid_cap(int64_t items, int per_block, int cap) {
    int64_t b = (items + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

template <int G>
void launch_propose(const DevGraph& g, CoarsenBufs& b, uint64_t seed, int shem, const int32_t* prev,
                    hipStream_t s) {
    constexpr int NG = kBlock / G;
    hipLaunchKernelGGL(k_match_propose<G>, dim3((g.n + NG - 1) / NG), dim3(kBlock), 0, s, g.n, g.xadj, g.adjncy,
                       g.adjwgt, b.match, b.cand, seed, shem, prev);
}

}  // namespace

void match_graph(const DevGraph& g, int ctype_shem, uint64_t seed, int rounds, CoarsenBufs& b, hipStream_t s) {
    if (g.n <= 0) return;
    if (rounds > kMaxMatchRounds) rounds = kMaxMatchRounds;
    const int nb = (g.n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_match_init, dim3(nb), dim3(kBlock), 0, s, g.n, g.xadj, b.match);
    RK_HIP(hipMemsetAsync(b.round_flags, 0, sizeof(int32_t) * kMaxMatchRounds, s));
    const int G = pick_group(g.n, g.m);
    for (int r = 0; r < rounds; ++r) {
        const int32_t* prev = r > 0 ? b.round_flags + (r - 1) : nullptr;
        switch (G) {
            case 4: launch_propose<4>(g, b, seed, ctype_shem, prev, s); break;
            case 8: launch_propose<8>(g, b, seed, ctype_shem, prev, s); break;
            case 16: launch_propose<16>(g, b, seed, ctype_shem, prev, s); break;
            case 32: launch_propose<32>(g, b, seed, ctype_shem, prev, s); break;
            default: launch_propose<64>(g, b, seed, ctype_shem, prev, s); break;
        }
        hipLaunchKernelGGL(k_match_accept, dim3(nb), dim3(kBlock), 0, s, g.n, b.cand, b.match,
                           b.round_flags + r, prev);
    }
    hipLaunchKernelGGL(k_match_finish, dim3(nb), dim3(kBlock), 0, s, g.n, b.match);
    RK_HIP(hipGetLastError());
}

int32_t contract_numbering(const DevGraph& g, const int32_t* match, int32_t* cmap, CoarsenBufs& b,
                           hipStream_t s) {
    (void)cmap;
    const int nb = (g.n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_leader, dim3(nb), dim3(kBlock), 0, s, g.n, match, b.leader);
    exclusive_scan_i32(b.leader, b.cid, g.n, b.counters + 0, b.scan_tmp, s);
    int32_t cn = 0;
    RK_HIP(hipMemcpyAsync(&cn, b.counters + 0, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RK_HIP(hipStreamSynchronize(s));
    return cn;
}

int32_t contract_build(const DevGraph& g, const int32_t* match, const int32_t* cmap, int32_t cn, DevGraph& cg,
                       bool aggregate_vsize, CoarsenBufs& b, CoarseAlloc& al, int* adjwgt_not_one, hipStream_t s) {
    const int nb = (g.n + kBlock - 1) / kBlock;
    cg.n = cn;
    cg.ncon = g.ncon;
    cg.vwgt = g.vwgt ? al.alloc_i32((int64_t)cn * g.ncon) : nullptr;
    cg.vsize = (aggregate_vsize && g.vsize) ? al.alloc_i32(cn) : nullptr;
    cg.xadj = al.alloc_i32((int64_t)cn + 1);
    hipLaunchKernelGGL(k_cmap, dim3(nb), dim3(kBlock), 0, s, g.n, g.ncon, g.xadj, g.vwgt, g.vsize, match, b.cid,
                       (int32_t*)cmap, b.members, b.degsum, cg.vwgt, cg.vsize);
    exclusive_scan_i32(b.degsum, b.tmp_off, cn, nullptr, b.scan_tmp, s);
    RK_HIP(hipMemsetAsync(b.counters + 2, 0, 6 * sizeof(int32_t), s));
    const int G = pick_group(g.n, g.m);
    const int NG = kBlock / G;
    const int grid = grid_cap(cn, NG, 65536);
#define RK_CONTRACT(GG)                                                                                     \
    hipLaunchKernelGGL(k_contract<GG>, dim3(grid), dim3(kBlock), 0, s, cn, g.xadj, g.adjncy, g.adjwgt, cmap, \
                       b.members, b.degsum, b.tmp_off, b.tmp_adj, b.tmp_w, b.cdeg, b.big_list, b.counters)
    switch (G) {
        case 4: RK_CONTRACT(4); break;
        case 8: RK_CONTRACT(8); break;
        case 16: RK_CONTRACT(16); break;
        case 32: RK_CONTRACT(32); break;
        default: RK_CONTRACT(64); break;
    }
#undef RK_CONTRACT
    hipLaunchKernelGGL(k_contract_big, dim3(1024), dim3(kBlock), 0, s, g.xadj, g.adjncy, g.adjwgt, cmap,
                       b.members, b.degsum, b.tmp_off, b.tmp_adj, b.tmp_w, b.cdeg, b.big_list, b.counters,
                       b.hash_keys, b.hash_first, b.hash_sum);
    exclusive_scan_i32(b.cdeg, cg.xadj, cn, cg.xadj + cn, b.scan_tmp, s);
    int32_t hc[2] = {0, 0};
    RK_HIP(hipMemcpyAsync(hc, cg.xadj + cn, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RK_HIP(hipMemcpyAsync(hc + 1, b.counters + 4, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RK_HIP(hipStreamSynchronize(s));
    const int32_t cm = hc[0];
    if (adjwgt_not_one) *adjwgt_not_one = hc[1];
    cg.m = cm;
    cg.adjncy = al.alloc_i32(cm > 0 ? cm : 1);
    cg.adjwgt = al.alloc_i32(cm > 0 ? cm : 1);
    hipLaunchKernelGGL(k_contract_copy<16>, dim3(grid_cap(cn, kBlock / 16, 65536)), dim3(kBlock), 0, s, cn,
                       b.tmp_off, b.tmp_adj, b.tmp_w, cg.xadj, cg.adjncy, cg.adjwgt);
    RK_HIP(hipGetLastError());
    return cm;
}

}  // namespace rk
