// coarsen.hip -- matching and coarse-graph contraction on the GPU.
//
// Matching (coarsen/shem.rs:13-56, coarsen/rm.rs:13-43).  The reference's
// matchings are sequential greedy scans.  On the GPU every unmatched vertex
// proposes to its best unmatched neighbour under a strict total order on
// edges (EdgeKey below) and mutual proposals become pairs (a
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
//    (323-341).  Coarse vertices with degree sum <= kContractBig are merged
//    in LDS tiles (k_contract_tile); larger ones use a block per vertex and
//    a hash table in global memory that records each coarse neighbour's
//    first position and weight sum (k_contract_big).
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
    // one plain store per block that changed something (every writer stores 1)
    if (__syncthreads_or(ch) && threadIdx.x == 0) *changed = 1;
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


constexpr int kCT = 256;                 // coarse vertices (= threads) per contraction tile
constexpr int kCItems = 16;              // staged positions per thread per batch
constexpr int kCCap = kCT * kCItems;     // merged-sequence positions staged per batch

__device__ inline int block_excl_scan_c(int x, int* s_w, int* total) {
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    int inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kCT / 64; ++w) {
        const int c = s_w[w];
        if (w < wave) off += c;
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return off + inc - x;
}

// One block per tile of 256 consecutive coarse vertices.  Their merged
// adjacency sequences adj(a) ++ adj(b) are streamed in batches of kCCap
// positions (independent coalesced loads of adjncy/adjwgt, then the cmap
// gathers) into LDS; each thread then merges its own coarse vertex in LDS:
// internal entries dropped, equal coarse neighbours summed onto their first
// occurrence, and the survivors written in sequence order.
__global__ __launch_bounds__(kCT) void k_contract_tile(const int32_t* __restrict__ cn_dev,
                                                       const int32_t* __restrict__ xadj,
                                                       const int32_t* __restrict__ adjncy,
                                                       const int32_t* __restrict__ adjwgt,
                                                       const int32_t* __restrict__ cmap,
                                                       const int32_t* __restrict__ members,
                                                       const int32_t* __restrict__ degsum,
                                                       const int32_t* __restrict__ tmp_off,
                                                       int32_t* __restrict__ tmp_adj, int32_t* __restrict__ tmp_w,
                                                       int32_t* __restrict__ cdeg, int32_t* __restrict__ big_list,
                                                       int32_t* __restrict__ counters) {
    __shared__ int32_t s_cu[kCCap];
    __shared__ int32_t s_w[kCCap];
    __shared__ int32_t s_loff[kCT + 1];
    __shared__ int32_t s_xa[kCT], s_xb[kCT], s_da[kCT];
    __shared__ int32_t s_bstart[kCT + 2];
    __shared__ int s_tmp[kCT / 64];
    __shared__ int s_nb;
    const int t = threadIdx.x;
    const int32_t cn = *cn_dev;
    if ((int64_t)blockIdx.x * kCT >= cn) return;   // block-uniform
    const int64_t cv = (int64_t)blockIdx.x * kCT + t;
    const bool own = cv < cn;
    int32_t ds = 0, xa = 0, xb = 0, da = 0;
    bool big = false;
    if (own) {
        const int32_t a = members[2 * cv], b = members[2 * cv + 1];
        ds = degsum[cv];
        xa = xadj[a];
        da = xadj[a + 1] - xa;
        xb = b >= 0 ? xadj[b] : 0;
        big = ds > kContractBig;
        if (big) big_list[atomicAdd(&counters[2], 1)] = (int32_t)cv;
    }
    int total;
    const int loff = block_excl_scan_c((own && !big) ? ds : 0, s_tmp, &total);
    s_loff[t] = loff;
    s_xa[t] = xa;
    s_xb[t] = xb;
    s_da[t] = da;
    if (t == 0) s_loff[kCT] = total;
    __syncthreads();
    if (t == 0) {
        int nb = 0, vs = 0;
        while (vs < kCT) {
            s_bstart[nb++] = vs;
            int lo = vs + 1, hi = kCT;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_loff[mid] - s_loff[vs] <= kCCap) lo = mid;
                else hi = mid - 1;
            }
            vs = lo;
        }
        s_bstart[nb] = kCT;
        s_nb = nb;
    }
    __syncthreads();
    bool not_one = false;
    const int nb = s_nb;
    for (int bb = 0; bb < nb; ++bb) {
        const int vs = s_bstart[bb], ve = s_bstart[bb + 1];
        const int P0 = s_loff[vs], P1 = s_loff[ve];
        int32_t uu[kCItems], ww[kCItems];
#pragma unroll
        for (int i = 0; i < kCItems; ++i) {
            const int p = P0 + t + i * kCT;
            uu[i] = 0;
            ww[i] = 1;
            if (p < P1) {
                int lo = vs, hi = ve - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_loff[mid] <= p) lo = mid;
                    else hi = mid - 1;
                }
                const int32_t tl = p - s_loff[lo];
                const int32_t j = tl < s_da[lo] ? s_xa[lo] + tl : s_xb[lo] + (tl - s_da[lo]);
                uu[i] = adjncy[j];
                if (adjwgt) ww[i] = adjwgt[j];
            }
        }
#pragma unroll
        for (int i = 0; i < kCItems; ++i) {
            const int p = P0 + t + i * kCT;
            if (p < P1) {
                s_cu[p - P0] = cmap[uu[i]];
                s_w[p - P0] = ww[i];
            }
        }
        __syncthreads();
        if (own && !big && t >= vs && t < ve) {
            const int lb = loff - P0, le = lb + ds;
            const int32_t me = (int32_t)cv;
            for (int e = lb; e < le; ++e)
                if (s_cu[e] == me) s_cu[e] = -1;                     // internal edge (graph/mod.rs:330-332)
            const int32_t off = tmp_off[cv];
            int32_t pos = 0;
            for (int e = lb; e < le; ++e) {
                const int32_t cu = s_cu[e];
                if (cu < 0) continue;
                uint32_t sum = (uint32_t)s_w[e];
                for (int e2 = e + 1; e2 < le; ++e2)
                    if (s_cu[e2] == cu) {                             // merge (graph/mod.rs:335-339)
                        sum += (uint32_t)s_w[e2];
                        s_cu[e2] = -1;
                    }
                tmp_adj[off + pos] = cu;
                tmp_w[off + pos] = (int32_t)sum;
                if ((int32_t)sum != 1) not_one = true;
                ++pos;
            }
            cdeg[cv] = pos;
        }
        __syncthreads();
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

__global__ __launch_bounds__(kBlock) void k_contract_copy(int32_t n_max, const int32_t* __restrict__ cn_dev,
                                                          const int32_t* __restrict__ tmp_off,
                                                          const int32_t* __restrict__ tmp_adj,
                                                          const int32_t* __restrict__ tmp_w,
                                                          const int32_t* __restrict__ cxadj,
                                                          int32_t* __restrict__ cadj, int32_t* __restrict__ cw) {
    constexpr int G = 16, NG = kBlock / G;
    const int gl = group_lane<G>();
    const int32_t cn = *cn_dev;
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

ContractResult contract(const DevGraph& g, const int32_t* match, int32_t* cmap, DevGraph& cg, bool aggregate_vsize,
                        CoarsenBufs& b, CoarseAlloc& al, int32_t* heavy_list, hipStream_t s) {
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
    hipLaunchKernelGGL(k_leader, dim3(nb), dim3(kBlock), 0, s, n, match, b.leader);
    exclusive_scan_i32(b.leader, b.cid, n, cn_dev, b.scan_tmp, s);      // coarse ids (graph/mod.rs:288-297)
    RK_HIP(hipMemsetAsync(b.degsum, 0, sizeof(int32_t) * n, s));
    RK_HIP(hipMemsetAsync(b.cdeg, 0, sizeof(int32_t) * n, s));
    RK_HIP(hipMemsetAsync(b.counters + 2, 0, 8 * sizeof(int32_t), s));
    hipLaunchKernelGGL(k_cmap, dim3(nb), dim3(kBlock), 0, s, n, g.ncon, g.xadj, g.vwgt, g.vsize, match, b.cid, cmap,
                       b.members, b.degsum, cg.vwgt, cg.vsize);
    exclusive_scan_i32(b.degsum, b.tmp_off, n, nullptr, b.scan_tmp, s);
    hipLaunchKernelGGL(k_contract_tile, dim3((n + kCT - 1) / kCT), dim3(kCT), 0, s, cn_dev, g.xadj, g.adjncy,
                       g.adjwgt, cmap, b.members, b.degsum, b.tmp_off, b.tmp_adj, b.tmp_w, b.cdeg, b.big_list,
                       b.counters);
    hipLaunchKernelGGL(k_contract_big, dim3(1024), dim3(kBlock), 0, s, g.xadj, g.adjncy, g.adjwgt, cmap,
                       b.members, b.degsum, b.tmp_off, b.tmp_adj, b.tmp_w, b.cdeg, b.big_list, b.counters,
                       b.hash_keys, b.hash_first, b.hash_sum);
    exclusive_scan_i32(b.cdeg, cg.xadj, n, cg.xadj + n, b.scan_tmp, s);   // cxadj[cn..n] = cm
    hipLaunchKernelGGL(k_contract_copy, dim3(grid_cap(n, kBlock / 16, 65536)), dim3(kBlock), 0, s, n, cn_dev,
                       b.tmp_off, b.tmp_adj, b.tmp_w, cg.xadj, cg.adjncy, cg.adjwgt);
    if (heavy_list) heavy_list_build(n, cn_dev, cg.xadj, heavy_list, b.counters + 8, s);
    int32_t hc[4] = {0, 0, 0, 0};
    RK_HIP(hipMemcpyAsync(hc + 0, cn_dev, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RK_HIP(hipMemcpyAsync(hc + 1, cg.xadj + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RK_HIP(hipMemcpyAsync(hc + 2, b.counters + 4, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RK_HIP(hipMemcpyAsync(hc + 3, b.counters + 8, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RK_HIP(hipStreamSynchronize(s));
    RK_HIP(hipGetLastError());
    cg.n = hc[0];
    cg.m = hc[1];
    ContractResult r;
    r.cn = hc[0];
    r.cm = hc[1];
    r.adjwgt_not_one = hc[2] != 0;
    r.nheavy = heavy_list ? hc[3] : 0;
    return r;
}

}  // namespace rk
