// grow.hip -- initial partition of the coarsest level by BFS growing
// (initial/grow.rs:12-103) on the device, one workgroup per graph.
//
// The reference grows the parts round-robin: in each round every part p
// that is below its targets takes vertices from the front of its BFS queue
// until one of them, v, still has unassigned neighbours; all of those become
// p's (in adjacency order, appended to p's queue), and only then does part
// p+1 take its turn.  Rounds repeat until none grows; the vertices left over
// go one by one, in vertex order, to the part lightest on constraint 0
// (first minimum).
//
// A round runs as sub-rounds over the parts not yet resolved in it (from f
// on).  Each such part proposes the first vertex of its queue with an
// unassigned neighbour (vertices with none are dropped for good: the
// unassigned set only shrinks) and stamps those neighbours with an atomic
// minimum of (sub-round, part): a neighbour goes to the lowest part that
// wants it, as in the reference's part order.  The parts before the first
// one, q, whose proposal wins no neighbour then take exactly the neighbours
// they won: every part before them took its proposal, so what a part finds
// unassigned in the reference is what no lower part claimed.  Part q's
// vertex lost every neighbour to lower parts, so the reference pops it
// without growth; it is dropped and the next sub-round restarts from q with
// the new assignments (q, now first, wins whatever it proposes).  Each
// sub-round resolves at least one part or drops one vertex.
#include <climits>

#include "common.hpp"
#include "kernels.hpp"

namespace rk {

namespace {

constexpr int kGrowT = 256;                      // threads per graph
constexpr uint64_t kStampTop = (1ull << 40) - 1;  // stamps count down from here, one per sub-round
constexpr int kGrowLdsParts = 4096;               // leftover weights of larger k stay in global memory
constexpr int kGrowClosedParts = 1024;            // the equal-weight leftover rule ranks at most this many parts
constexpr int kShortRow = 96;                     // rows longer than this are scanned by a whole wave
constexpr int kGrowMaxCoopCon = 4;                // constraints summed in registers by a wave

__device__ inline int32_t vw_at(const int32_t* vwgt, int64_t gv, int32_t ncon, int32_t c) {
    return vwgt ? vwgt[gv * ncon + c] : 1;
}

// 64-bit minimum over the wave by DPP row shifts and broadcasts (the
// result of lane 63, returned to every lane).
template <int CTRL, int ROWS>
__device__ inline uint64_t dpp_min_step(uint64_t x) {
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    const uint32_t ylo = (uint32_t)__builtin_amdgcn_update_dpp((int)lo, (int)lo, CTRL, ROWS, 0xF, false);
    const uint32_t yhi = (uint32_t)__builtin_amdgcn_update_dpp((int)hi, (int)hi, CTRL, ROWS, 0xF, false);
    const uint64_t y = ((uint64_t)yhi << 32) | ylo;
    return y < x ? y : x;
}

__device__ inline uint64_t wave_min_u64(uint64_t x) {
    x = dpp_min_step<0x111, 0xF>(x);    // row_shr:1
    x = dpp_min_step<0x112, 0xF>(x);    // row_shr:2
    x = dpp_min_step<0x114, 0xF>(x);    // row_shr:4
    x = dpp_min_step<0x118, 0xF>(x);    // row_shr:8: lane 15 of each row holds the row's minimum
    x = dpp_min_step<0x142, 0xA>(x);    // row_bcast:15 into rows 1 and 3
    x = dpp_min_step<0x143, 0xC>(x);    // row_bcast:31 into rows 2 and 3: lane 63 holds the minimum
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), 63);
    return ((uint64_t)hi << 32) | lo;
}

__device__ inline long long block_min_ll(long long x, long long* s_red) {
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = min(x, (long long)__shfl_xor(x, o, 64));
    if (lane == 0) s_red[wave] = x;
    __syncthreads();
    long long r = s_red[0];
#pragma unroll
    for (int w = 1; w < kGrowT / 64; ++w) r = min(r, s_red[w]);
    __syncthreads();
    return r;
}

__device__ inline long long block_sum_ll(long long x, long long* s_red) {
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += (long long)__shfl_xor(x, o, 64);
    if (lane == 0) s_red[wave] = x;
    __syncthreads();
    long long r = 0;
#pragma unroll
    for (int w = 0; w < kGrowT / 64; ++w) r += s_red[w];
    __syncthreads();
    return r;
}

__device__ inline int block_min_int(int x, int* s_red) {
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o, 64));
    if (lane == 0) s_red[wave] = x;
    __syncthreads();
    int r = s_red[0];
#pragma unroll
    for (int w = 1; w < kGrowT / 64; ++w) r = min(r, s_red[w]);
    __syncthreads();
    return r;
}

// One graph.  LDS: the graph's CSR (rebased to local ids) and every table
// live in the workgroup's LDS (a graph small enough, see grow_lds_bytes);
// otherwise the tables are the global scratch of ga and the CSR is read in
// place.  The result goes to ga.out (graph-local).
template <bool LDS>
__device__ __forceinline__ void grow_body(const GrowArgs& ga, char* s_dyn, int* s_red, long long* s_red64,
                                          int* s_cnt) {
    const int32_t* __restrict__ xadj = ga.xadj;
    const int32_t* __restrict__ adjncy = ga.adjncy;
    const int32_t* __restrict__ vwgt = ga.vwgt;
    const int tid = threadIdx.x;
    const int32_t n = ga.n, k = ga.k, ncon = ga.ncon;
    const int64_t v0 = ga.v0, e0 = ga.e0;
    int32_t *P, *qnext, *left, *st, *X = nullptr, *A = nullptr;
    unsigned long long* owner;
    long long *pw, *W;
    if constexpr (LDS) {
        owner = (unsigned long long*)s_dyn;                       // grow_lds_bytes' layout
        pw = (long long*)(owner + n);
        W = pw + (int64_t)k * ncon;
        P = (int32_t*)(W + k);
        qnext = P + n;
        left = qnext + n;
        X = left + n;
        A = X + n + 1;
        st = A + ga.m;
        for (int32_t v = tid; v <= n; v += kGrowT) X[v] = xadj[v0 + v] - (int32_t)e0;
        for (int32_t j = tid; j < ga.m; j += kGrowT) A[j] = adjncy[e0 + j] - (int32_t)v0;
    } else {
        owner = ga.owner;
        pw = ga.pw;
        W = k <= kGrowLdsParts ? (long long*)s_dyn : ga.w0;
        P = ga.out;
        qnext = ga.qnext;
        left = ga.leftover;
        st = ga.st;
    }
    int32_t *head = st, *tail = st + k, *prop = st + 2 * (int64_t)k, *act = st + 3 * (int64_t)k;
    // row of local vertex v: [rb(v), rb(v+1)); neighbour of entry j: nb(j) (local ids)
    auto rb = [&](int32_t v) __attribute__((always_inline)) -> int64_t {
        if constexpr (LDS) return X[v];
        else return xadj[v0 + v];
    };
    auto nb = [&](int64_t j) __attribute__((always_inline)) -> int32_t {
        if constexpr (LDS) return A[j];
        else return adjncy[j] - (int32_t)v0;
    };
    int32_t *listA = st + 4 * (int64_t)k, *listB = st + 5 * (int64_t)k, *listC = st + 6 * (int64_t)k;
    int32_t* first = left;                        // first position of a claimed vertex in a long row (wave scans)
    // row scans with 8 loads in flight
    auto row_has_free = [&](int32_t v) __attribute__((always_inline)) -> bool {
        for (int64_t j = rb(v), e = rb(v + 1); j < e; j += 8) {
            int32_t q[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) q[t] = j + t < e ? P[nb(j + t)] : 0;
            for (int t = 0; t < 8; ++t)
                if (q[t] == -1) return true;
        }
        return false;
    };
    auto row_wins = [&](int32_t v, unsigned long long mine) __attribute__((always_inline)) -> bool {
        for (int64_t j = rb(v), e = rb(v + 1); j < e; j += 8) {
            unsigned long long o[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) o[t] = j + t < e ? owner[nb(j + t)] : 0ull;
            for (int t = 0; t < 8; ++t)
                if (o[t] == mine) return true;
        }
        return false;
    };
    // One wave on one long row (called by every lane of the wave; v, p
    // wave-uniform): the waves of the block take the listed rows in turn.
    const int wave = tid >> 6, lane = __lane_id();
    constexpr int kWaves = kGrowT / 64;
    auto co_has_free = [&](int32_t v) __attribute__((always_inline)) -> bool {
        for (int64_t c0 = rb(v), e = rb(v + 1); c0 < e; c0 += 64) {
            const int64_t j = c0 + lane;
            if (__ballot(j < e && P[nb(j)] == -1)) return true;
        }
        return false;
    };
    auto co_claim = [&](int32_t v, unsigned long long mine) __attribute__((always_inline)) {
        for (int64_t j = rb(v) + lane, e = rb(v + 1); j < e; j += 64) {
            const int32_t u = nb(j);
            if (P[u] == -1) atomicMin(&owner[u], mine);
        }
    };
    auto co_wins = [&](int32_t v, unsigned long long mine) __attribute__((always_inline)) -> bool {
        for (int64_t c0 = rb(v), e = rb(v + 1); c0 < e; c0 += 64) {
            const int64_t j = c0 + lane;
            if (__ballot(j < e && owner[nb(j)] == mine)) return true;
        }
        return false;
    };
    // takes v's won neighbours for p in row order: per 64 entries of the row
    // the claimers (a repeated entry only at its first position) are linked
    // after the queue's tail in lane order
    auto co_commit = [&](int32_t p, int32_t v, unsigned long long mine) __attribute__((always_inline)) {
        int32_t h = qnext[v], t = h < 0 ? -1 : tail[p];   // pop v
        long long wsum[kGrowMaxCoopCon];
#pragma unroll
        for (int c = 0; c < kGrowMaxCoopCon; ++c) wsum[c] = 0;
        const bool regs = ncon <= kGrowMaxCoopCon;    // else the weights go to pw by atomics
        const int64_t r0 = rb(v), e = rb(v + 1);
        for (int64_t c0 = r0; c0 < e; c0 += 64) {
            const int64_t j = c0 + lane;
            int32_t u = -1;
            bool cl = false;
            if (j < e) {
                u = nb(j);
                cl = owner[u] == mine && P[u] == -1;
            }
            if (!ga.norep) {                    // a repeated entry only at its first position
                if (cl) atomicMin(&first[u], (int32_t)(j - r0));
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                if (cl) cl = __hip_atomic_load(&first[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == (int32_t)(j - r0);
            }
            const uint64_t bm = __ballot(cl);
            if (bm == 0) continue;
            const uint64_t below = bm & ((1ull << lane) - 1ull);
            const int32_t pu = __shfl(u, below ? 63 - __builtin_clzll(below) : 0, 64);
            if (cl) {
                P[u] = p;
                const int32_t prev = below ? pu : t;
                if (prev >= 0) qnext[prev] = u;
                if (regs) {
#pragma unroll
                    for (int c = 0; c < kGrowMaxCoopCon; ++c)
                        if (c < ncon) wsum[c] += vw_at(vwgt, v0 + u, ncon, c);
                } else {
                    for (int32_t c = 0; c < ncon; ++c)
                        atomicAdd((unsigned long long*)&pw[(int64_t)p * ncon + c],
                                  (unsigned long long)vw_at(vwgt, v0 + u, ncon, c));
                }
            }
            if (h < 0) h = __shfl(u, __builtin_ctzll(bm), 64);
            t = __shfl(u, 63 - __builtin_clzll(bm), 64);
        }
#pragma unroll
        for (int c = 0; c < kGrowMaxCoopCon; ++c) {
            if (!regs || c >= ncon) break;
            long long x = wsum[c];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) x += (long long)__shfl_xor(x, o, 64);
            if (lane == 0) pw[(int64_t)p * ncon + c] += x;
        }
        if (lane == 0) {
            head[p] = h;
            tail[p] = t;
        }
    };
    for (int32_t v = tid; v < n; v += kGrowT) {
        P[v] = -1;
        qnext[v] = -1;
        owner[v] = ~0ull;
        first[v] = INT_MAX;
    }
    for (int64_t i = tid; i < (int64_t)k * ncon; i += kGrowT) pw[i] = 0;
    __syncthreads();
    for (int32_t p = tid; p < k; p += kGrowT) {     // grow.rs:41-47 (the seeds are distinct)
        const int32_t sd = ga.seeds[p];
        P[sd] = p;
        head[p] = sd;
        tail[p] = sd;
        for (int32_t c = 0; c < ncon; ++c) pw[(int64_t)p * ncon + c] = vw_at(vwgt, v0 + sd, ncon, c);
    }
    __syncthreads();
    uint64_t sub = 0;
    for (;;) {                                      // grow.rs:50-86, one round per iteration
        for (int32_t p = tid; p < k; p += kGrowT) {
            bool at = true;                         // grow.rs:55-57
            for (int32_t c = 0; c < ncon; ++c)
                if (!((double)pw[(int64_t)p * ncon + c] >= ga.tpw[(int64_t)p * ncon + c])) at = false;
            act[p] = (!at && head[p] >= 0) ? 1 : 0;
        }
        if (tid == 0) s_cnt[0] = 0;                 // parts grown in the round
        __syncthreads();
        for (int32_t f = 0;;) {                     // sub-rounds; f is block-uniform
            ++sub;
            const unsigned long long stamp = (unsigned long long)(kStampTop - sub) << 24;
            if (tid == 0) s_cnt[2] = s_cnt[3] = s_cnt[4] = 0;   // long-row lists of the three phases
            __syncthreads();
            // proposals (rows longer than kShortRow are listed for the block)
            for (int32_t p = f + tid; p < k; p += kGrowT) {
                prop[p] = -1;
                if (!act[p]) continue;
                int32_t v = head[p];
                bool lng = false;
                for (; v >= 0; v = qnext[v]) {      // drop queue vertices without unassigned neighbours
                    if (rb(v + 1) - rb(v) > kShortRow) {
                        lng = true;
                        break;
                    }
                    if (row_has_free(v)) break;
                }
                head[p] = v;
                if (v < 0) {                        // queue empty: no growth this round (or later)
                    act[p] = 0;
                    continue;
                }
                if (lng) {
                    listA[atomicAdd(&s_cnt[2], 1)] = p;
                    continue;
                }
                prop[p] = v;
                for (int64_t j = rb(v), e = rb(v + 1); j < e; ++j) {
                    const int32_t u = nb(j);
                    if (P[u] == -1) atomicMin(&owner[u], stamp | (unsigned long long)p);
                }
            }
            __syncthreads();
            for (int32_t i = wave, nd = s_cnt[2]; i < nd; i += kWaves) {
                const int32_t p = listA[i];
                int32_t v = head[p];
                while (v >= 0 && !co_has_free(v)) v = qnext[v];
                if (v >= 0) co_claim(v, stamp | (unsigned long long)p);
                if (lane == 0) {
                    head[p] = v;
                    if (v < 0) act[p] = 0;
                    else prop[p] = v;
                }
            }
            __syncthreads();
            // the first part (>= f) whose proposal wins no neighbour; each
            // thread's parts ascend, so its first such part among those it
            // checks is its minimum (long rows are checked by the block)
            int fm = k;
            for (int32_t p = f + tid; p < k; p += kGrowT) {
                const int32_t v = prop[p];
                if (v < 0) continue;
                if (rb(v + 1) - rb(v) > kShortRow) {
                    listB[atomicAdd(&s_cnt[3], 1)] = p;
                    continue;
                }
                const bool win = row_wins(v, stamp | (unsigned long long)p);
                if (!win) {
                    fm = p;
                    break;
                }
            }
            __syncthreads();
            {
                int fw = k;                         // wave-uniform: the waves' failing long rows
                for (int32_t i = wave, nd = s_cnt[3]; i < nd; i += kWaves) {
                    const int32_t p = listB[i];
                    if (p < fw && !co_wins(prop[p], stamp | (unsigned long long)p)) fw = p;
                }
                fm = min(fm, fw);
            }
            fm = block_min_int(fm, s_red);
            // parts [f, fm) take what they won (grow.rs:62-78); their won sets
            // are disjoint, so their writes never meet
            int grew = 0;
            for (int32_t p = f + tid; p < fm; p += kGrowT) {
                const int32_t v = prop[p];
                if (v < 0) continue;
                if (rb(v + 1) - rb(v) > kShortRow) {
                    listC[atomicAdd(&s_cnt[4], 1)] = p;
                    continue;
                }
                int32_t h = qnext[v], t = h < 0 ? -1 : tail[p];   // pop v
                for (int64_t j = rb(v), e = rb(v + 1); j < e; ++j) {
                    const int32_t u = nb(j);
                    // the ownership test first (only p writes the parts of
                    // what it owns); the part test skips a repeated entry
                    if (owner[u] != (stamp | (unsigned long long)p) || P[u] != -1) continue;
                    P[u] = p;
                    for (int32_t c = 0; c < ncon; ++c) pw[(int64_t)p * ncon + c] += vw_at(vwgt, v0 + u, ncon, c);
                        if (h < 0) h = u;
                    else qnext[t] = u;
                    t = u;
                }
                head[p] = h;
                tail[p] = t;
                ++grew;
            }
            if (grew) atomicAdd(&s_cnt[0], grew);
            __syncthreads();
            for (int32_t i = wave, nd = s_cnt[4]; i < nd; i += kWaves) {
                const int32_t p = listC[i];
                co_commit(p, prop[p], stamp | (unsigned long long)p);
                if (lane == 0) atomicAdd(&s_cnt[0], 1);
            }
            if (fm < k && tid == 0) head[fm] = qnext[prop[fm]];   // fm's vertex lost every neighbour: popped
            __syncthreads();
            if (fm >= k) break;
            f = fm;
        }
        __syncthreads();
        const int grown = s_cnt[0];
        __syncthreads();
        if (grown == 0) break;                      // grow.rs:50: no part grew in the round
    }
    // leftovers (grow.rs:89-100), in vertex order, to the lightest part on
    // constraint 0 (first minimum): listed in order, then walked by wave 0
    // with each lane holding the minimum of its parts (p = lane + 64 j)
    if (tid == 0) s_cnt[1] = 0;
    __syncthreads();
    for (int32_t base = 0; base < n; base += kGrowT) {
        const int32_t v = base + tid;
        const bool lv = v < n && P[v] == -1;
        const uint64_t m = __ballot(lv);
        // ordered compaction: wave offsets from the waves' popcounts
        if (__lane_id() == 0) s_red[threadIdx.x >> 6] = __popcll(m);
        __syncthreads();
        int off = s_cnt[1];
        for (int w = 0; w < (threadIdx.x >> 6); ++w) off += s_red[w];
        if (lv) left[off + __popcll(m & ((1ull << __lane_id()) - 1ull))] = v;
        __syncthreads();
        if (tid == 0) {
            int tot = 0;
            for (int w = 0; w < kGrowT / 64; ++w) tot += s_red[w];
            s_cnt[1] += tot;
        }
        __syncthreads();
    }
    const int32_t nleft = s_cnt[1];
    // Equal leftover weights w: the walk's choices are the slots W_p + j w
    // (p's weight before the leftovers, j >= 0) in (value, part) order.  With
    // base the lightest weight, band t = [base + t w, base + (t+1) w) holds
    // one slot of every part p whose band t_p = (W_p - base) / w is t or
    // below, in order of (r_p, p), r_p = (W_p - base) mod w.  Parts join at
    // their band; between two joins the bands list the same parts, so the
    // leftovers of such a segment of bands go to its list in turn.
    bool closed = false;
    if (nleft > 0 && k <= kGrowClosedParts) {
        int wmin = INT_MAX, wmax = 0;
        for (int32_t i = tid; i < nleft; i += kGrowT) {
            const int w = vw_at(vwgt, v0 + left[i], ncon, 0);
            wmin = min(wmin, w);
            wmax = max(wmax, w);
        }
        wmin = block_min_int(wmin, s_red);
        wmax = -block_min_int(-wmax, s_red);
        closed = wmin == wmax;
        if (closed) {
            // per-part tables in the (now free) queue tables: band, residue,
            // rank in (r, p) order and in (t, p) order, the join order, a
            // segment's list and the (r, p) order
            int32_t *tb = st, *rs = st + k, *rrank = st + 2 * (int64_t)k, *trank = st + 3 * (int64_t)k;
            int32_t *join = st + 4 * (int64_t)k, *lst = st + 5 * (int64_t)k, *rord = st + 6 * (int64_t)k;
            long long base = LLONG_MAX;
            for (int32_t p = tid; p < k; p += kGrowT) base = min(base, pw[(int64_t)p * ncon]);
            base = block_min_ll(base, s_red64);
            for (int32_t p = tid; p < k; p += kGrowT) {
                const long long d = pw[(int64_t)p * ncon] - base;
                tb[p] = (int32_t)(d / wmin);
                rs[p] = (int32_t)(d % wmin);
            }
            __syncthreads();
            for (int32_t p = tid; p < k; p += kGrowT) {
                int32_t a1 = 0, a2 = 0;
                for (int32_t q = 0; q < k; ++q) {
                    a1 += (rs[q] < rs[p] || (rs[q] == rs[p] && q < p)) ? 1 : 0;
                    a2 += (tb[q] < tb[p] || (tb[q] == tb[p] && q < p)) ? 1 : 0;
                }
                rrank[p] = a1;
                trank[p] = a2;
            }
            __syncthreads();
            for (int32_t p = tid; p < k; p += kGrowT) {
                join[trank[p]] = p;
                rord[rrank[p]] = p;
            }
            __syncthreads();
            int64_t start = 0;
            for (int32_t jn = 0; start < nleft;) {       // segments; block-uniform
                const int32_t tcur = tb[join[jn]];
                int32_t je = jn;
                while (je < k && tb[join[je]] == tcur) ++je;   // the parts joining at tcur
                const int64_t seg = je < k ? (int64_t)je * (int64_t)(tb[join[je]] - tcur) : INT64_MAX;
                // the joined parts (t_p <= tcur) in (r, p) order: an ordered
                // compaction of the (r, p) order rord
                int32_t cnt = 0;                            // block-uniform
                for (int32_t c0 = 0; c0 < k; c0 += kGrowT) {
                    const int32_t q = c0 + tid;
                    const bool in = q < k && tb[rord[q]] <= tcur;
                    const uint64_t bm = __ballot(in);
                    if (lane == 0) s_red[wave] = __popcll(bm);
                    __syncthreads();
                    int off = cnt, tot = 0;
                    for (int w = 0; w < kWaves; ++w) {
                        off += w < wave ? s_red[w] : 0;
                        tot += s_red[w];
                    }
                    if (in) lst[off + __popcll(bm & ((1ull << lane) - 1ull))] = rord[q];
                    cnt += tot;
                    __syncthreads();
                }
                const int64_t end = seg == INT64_MAX || start + seg > nleft ? nleft : start + seg;
                for (int64_t i = start + tid; i < end; i += kGrowT) P[left[i]] = lst[(i - start) % je];
                __syncthreads();
                start = end;
                jn = je;
            }
        }
    }
    const int64_t walk = closed ? 0 : nleft;
    if (walk > 0 && tid < 64) {
        // wave 0 walks the leftovers; lane l keeps the smallest
        // key (weight << 24 | part) of its parts p = l + 64 j, and each W[p]
        // is read and written by lane p % 64 only
        const int lane = tid;
        for (int32_t p = lane; p < k; p += 64) W[p] = pw[(int64_t)p * ncon];
        uint64_t lk = ~0ull;
        auto lane_min = [&]() __attribute__((always_inline)) {
            lk = ~0ull;
            for (int32_t p = lane; p < k; p += 64) lk = min(lk, ((uint64_t)W[p] << 24) | (uint64_t)p);
        };
        lane_min();
        for (int64_t i = 0; i < walk; ++i) {
            const int32_t v = left[i];
            const int32_t mp = (int32_t)(wave_min_u64(lk) & 0xFFFFFFu);   // first minimum: ties by id
            if (lane == 0) P[v] = mp;
            if (lane == (mp & 63)) {
                W[mp] += vw_at(vwgt, v0 + v, ncon, 0);
                lane_min();
            }
        }
    }
    if constexpr (LDS) {
        __syncthreads();
        for (int32_t v = tid; v < n; v += kGrowT) ga.out[v] = P[v];
    }
}

__global__ __launch_bounds__(kGrowT) void k_grow(const GrowArgs* __restrict__ table) {
    extern __shared__ __align__(16) char s_dyn[];
    __shared__ int s_red[kGrowT / 64];
    __shared__ long long s_red64[kGrowT / 64];
    __shared__ int s_cnt[5];
    const GrowArgs ga = table[blockIdx.x];
    if (ga.lds) grow_body<true>(ga, s_dyn, s_red, s_red64, s_cnt);
    else grow_body<false>(ga, s_dyn, s_red, s_red64, s_cnt);
}

}  // namespace

size_t grow_lds_bytes(int32_t n, int64_t m, int32_t k, int32_t ncon) {
    // owner, part weights, leftover weights (8 B each); parts, queue links,
    // leftovers, the rebased CSR and the seven per-part tables (4 B each)
    return 8 * ((size_t)n + (size_t)k * ncon + (size_t)k) + 4 * (3 * (size_t)n + (size_t)n + 1 + (size_t)m + 7 * (size_t)k);
}

size_t grow_global_lds_bytes(int32_t k) { return sizeof(long long) * (size_t)std::min(k, kGrowLdsParts); }

void grow_partitions(const GrowArgs* table, int32_t ngraphs, size_t lds, hipStream_t s) {
    if (ngraphs <= 0) return;
    hipLaunchKernelGGL(k_grow, dim3(ngraphs), dim3(kGrowT), lds, s, table);
    RK_HIP(hipGetLastError());
}

}  // namespace rk
