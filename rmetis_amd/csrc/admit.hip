// admit.hip -- ordered admission of one greedy pass (refine/greedy.rs:64-123,
// 131-181) in a single workgroup.
//
// The reference visits the candidate movers in ascending vertex order; each
// one takes the first part of its gain-ordered candidate list that is
// feasible under the *running* part weights (and, in the repair pass, only if
// its own part is still overweight), then updates those weights.  Passes
// carry tens to a few thousand candidates, and every decision may depend on
// the one before it, so this step is a latency-bound sequential walk:
//
//  * all 256 threads list the tiles with movers (a bitmap set by the gain
//    kernel), gather their records into one dense ordered list, and stage
//    windows of 256 movers (record, vertex weights, first candidates) in LDS;
//  * the running state is one LDS table slack[p][c] = fmax[p][c] - pw[p][c]:
//    part p can take a vertex of weight w iff w <= slack[p][c] for every c;
//  * wave 0 walks the window one mover at a time: lane t reads the slack of
//    the mover's t-th candidate (one LDS round trip for the whole list), a
//    ballot picks the first feasible one, and the lanes of the constraints
//    apply the move with LDS atomics.  LDS instructions of one wave execute
//    in issue order, so the next mover's reads see those updates;
//  * the window's decisions are then written out by all threads (moves in
//    admission order, the new part of each moved vertex).
//
// Thresholds floor(tpw*ub) are clamped to [-1, 2^31-1]: with positive vertex
// weights and every per-constraint total below 2^31 (checked by the callers)
// each running weight and weight+vw lies in [0, 2^31), so the clamped
// compares equal the reference's f64 compares (greedy.rs:70,87,144) and the
// slack always fits in int32.
#include <climits>
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace rk {

namespace {

constexpr int kAT = 256;        // threads per admission workgroup
constexpr int kLdsKC = 4096;    // k*ncon entries of slack/thresholds kept in LDS
constexpr int kCL = 8;          // candidates per mover staged in LDS
constexpr int kWin = 256;       // movers staged per window
constexpr int kCntWin = 2048;   // tiles with movers gathered per window
constexpr int kNoDest = -1;
constexpr size_t kLdsBytes = 64 * 1024;   // dynamic LDS of one admission workgroup


// LDS layout of the mover window.
struct Win {
    int4* rec;       // [kWin]  (v, pv, ncand, nbefore | zero-group flag << 31)
    int32_t* off;    // [kWin]  xadj[v]
    int32_t* w0;     // [kWin]
    int32_t* w1;     // [kWin]
    int32_t* dst;    // [kWin]  decision per slot (-1: stays)
    int32_t* cl;     // [kWin * kCL]
};
constexpr int kWinInts = kWin * (4 + 1 + 1 + 1 + 1 + kCL);

#define RK_INL __device__ __attribute__((always_inline)) inline

RK_INL int32_t clamp_thr(long long x) {
    return x > INT32_MAX ? INT32_MAX : (x < -1 ? -1 : (int32_t)x);
}

RK_INL int block_excl_scan_int(int x, int* s_w, int* total) {
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
    for (int w = 0; w < kAT / 64; ++w) {
        const int c = s_w[w];
        if (w < wave) off += c;
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return off + inc - x;
}

// Tables of the pass.  Every helper is force-inlined so that, in the
// LDS-resident instantiation, these pointers keep the LDS address space
// (ds_* accesses waited on with lgkmcnt only).
struct Ctx {
    const AdmitArgs* a;
    int32_t* slack;       // fmax - pw, running
    const int32_t* fm;    // clamped fmax
    const int32_t* om;    // clamped omax
    Win w;
};

// Weight of mover v in constraint c; constraints 0 and 1 come staged.
RK_INL int32_t vw_c(const AdmitArgs& a, int32_t v, int c, int32_t w0, int32_t w1) {
    return c == 0 ? w0 : (c == 1 ? w1 : a.vwgt[(int64_t)v * a.ncon + c]);
}

// Candidate t of the mover in window slot s.  The staged copy is always read
// (at an in-range index) and the pool only beyond it, so the select is on
// values and the LDS read stays an LDS read.
RK_INL int32_t cand_at(const Ctx& x, int s, int32_t off, int t) {
    const int32_t staged = x.w.cl[s * kCL + (t & (kCL - 1))];
    const int32_t pooled = t >= kCL ? x.a->pool[off + t] : 0;
    return t < kCL ? staged : pooled;
}

// Can part q take mover v (weights w0, w1, ...) under the running slack?
RK_INL bool fits(const Ctx& x, int32_t q, int32_t v, int32_t w0, int32_t w1) {
    const int ncon = x.a->ncon;
    if (ncon == 1) return w0 <= x.slack[q];
    bool ok = w0 <= x.slack[q * ncon] && w1 <= x.slack[q * ncon + 1];
    for (int c = 2; c < ncon; ++c) ok = ok && vw_c(*x.a, v, c, w0, w1) <= x.slack[q * ncon + c];
    return ok;
}

RK_INL uint64_t low_mask(int n) { return n >= 64 ? ~0ull : ((1ull << n) - 1ull); }

// Exact decision of one mover against the running weights.  The lanes test
// candidate parts in parallel and the first feasible one in admission order
// wins: the explicit candidates with gain above the zero group, then the
// zero group (parts with no edge to v, ascending id), then the rest of the
// explicit list.  All arguments are wave-uniform.
// q0 holds the mover's first 64 candidates, one per lane (-1 past the list).
RK_INL int decide_table(const Ctx& x, int s, int32_t v, int32_t pv, int32_t nc, int32_t nbz, int32_t off,
                        int32_t w0, int32_t w1, int32_t q0) {
    const AdmitArgs& a = *x.a;
    const int lane = __lane_id();
    const int ncon = a.ncon;
    if (a.mode == 0) {                                               // greedy.rs:69-74
        bool over = false;
        for (int c = lane; c < ncon; c += 64)
            over = over ||
                   (long long)x.fm[pv * ncon + c] - x.slack[pv * ncon + c] > (long long)x.om[pv * ncon + c];
        if (__ballot(over) == 0) return kNoDest;
    }
    const int nbefore = nbz & 0x7fffffff;
    const bool zflag = ((uint32_t)nbz >> 31) != 0;
    const uint64_t f0 = __ballot(q0 >= 0 && fits(x, q0, v, w0, w1));
    // explicit candidates [0, nb): the whole list, or those ahead of the zero group
    const int nb = zflag ? nbefore : nc;
    if (const uint64_t m = f0 & low_mask(nb)) return __builtin_amdgcn_readlane(q0, __builtin_ctzll(m));
    for (int tb = 64; tb < nb; tb += 64) {
        const int t = tb + lane;
        const int32_t q = t < nc ? cand_at(x, s, off, t) : -1;
        if (const uint64_t m = __ballot(t < nb && q >= 0 && fits(x, q, v, w0, w1)))
            return __builtin_amdgcn_readlane(q, __builtin_ctzll(m));
    }
    if (!zflag) return kNoDest;
    // zero group: parts other than pv that are not in the explicit list
    for (int qb = 0; qb < a.k; qb += 64) {
        const int32_t q = qb + lane;
        bool listed = q == pv;
        for (int t = 0; t < nc; ++t) {
            const int32_t ct = t < 64 ? __builtin_amdgcn_readlane(q0, t) : cand_at(x, s, off, t);
            listed = listed || ct == q;
        }
        if (const uint64_t m = __ballot(q < a.k && !listed && fits(x, q, v, w0, w1))) return qb + __builtin_ctzll(m);
    }
    // explicit candidates [nbefore, nc)
    if (const uint64_t m = f0 & ~low_mask(nbefore)) return __builtin_amdgcn_readlane(q0, __builtin_ctzll(m));
    for (int tb = 64; tb < nc; tb += 64) {
        if (tb + 64 <= nbefore) continue;
        const int t = tb + lane;
        const int32_t q = t < nc ? cand_at(x, s, off, t) : -1;
        if (const uint64_t m = __ballot(t >= nbefore && q >= 0 && fits(x, q, v, w0, w1)))
            return __builtin_amdgcn_readlane(q, __builtin_ctzll(m));
    }
    return kNoDest;
}

// Running slack in the LDS (or, for large k*ncon, global) table.  LDS
// operations of one wave are issued and completed in order, so the next
// mover's reads see the atomics of this one; in global memory a fence orders
// them.
template <bool L>
struct TableSlack {
    RK_INL int decide(const Ctx& x, int s, int32_t v, int32_t pv, int32_t nc, int32_t nbz, int32_t off, int32_t w0,
                      int32_t w1, int32_t q0) const {
        return decide_table(x, s, v, pv, nc, nbz, off, w0, w1, q0);
    }
    RK_INL void apply(const Ctx& x, int32_t v, int32_t d, int32_t pv, int32_t w0, int32_t w1) {
        const int ncon = x.a->ncon;
        for (int c = __lane_id(); c < ncon; c += 64) {   // one constraint per lane
            const int32_t wc = vw_c(*x.a, v, c, w0, w1);
            atomicSub(&x.slack[d * ncon + c], wc);
            atomicAdd(&x.slack[pv * ncon + c], wc);
        }
        if constexpr (!L) __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    }
    RK_INL void load(const Ctx&) {}
    RK_INL void store(const Ctx&) const {}
};

// Running slack in wave 0's registers for k <= 64*KP and ncon == NC <= 2:
// lane l of s0[j] / s1[j] holds part 64*j + l (INT32_MIN past k, never
// feasible).  A mover's feasible parts are one compare per register block
// and constraint; no memory access sits on the mover-to-mover chain.
template <int KP, int NC>
struct RegSlack {
    int32_t s0[KP], s1[KP];

    RK_INL void load(const Ctx& x) {
        const int lane = __lane_id(), k = x.a->k;
#pragma unroll
        for (int j = 0; j < KP; ++j) {
            const int p = 64 * j + lane;
            s0[j] = p < k ? x.slack[p * NC] : INT32_MIN;
            s1[j] = p < k && NC > 1 ? x.slack[p * NC + 1] : INT32_MIN;
        }
    }
    RK_INL void store(const Ctx& x) const {
        const int lane = __lane_id(), k = x.a->k;
#pragma unroll
        for (int j = 0; j < KP; ++j) {
            const int p = 64 * j + lane;
            if (p < k) {
                x.slack[p * NC] = s0[j];
                if (NC > 1) x.slack[p * NC + 1] = s1[j];
            }
        }
    }
    // uniform part p: its slack in constraint c
    RK_INL int32_t at(int32_t p, int c) const {
        int32_t r = c == 0 ? s0[0] : s1[0];
#pragma unroll
        for (int j = 1; j < KP; ++j)
            if ((p >> 6) == j) r = c == 0 ? s0[j] : s1[j];
        return __builtin_amdgcn_readlane(r, p & 63);
    }
    // The decision walks the candidate list in admission order with scalar
    // code: F holds one feasibility bit per part, so each candidate costs a
    // readlane and a bit test, and most movers stop at the first one.
    RK_INL int decide(const Ctx& x, int s, int32_t, int32_t pv, int32_t nc, int32_t nbz, int32_t off, int32_t w0,
                      int32_t w1, int32_t q0) const {
        const AdmitArgs& a = *x.a;
        if (a.mode == 0) {                                           // greedy.rs:69-74
            bool over = (long long)x.fm[pv * NC] - at(pv, 0) > (long long)x.om[pv * NC];
            if (NC > 1) over = over || (long long)x.fm[pv * NC + 1] - at(pv, 1) > (long long)x.om[pv * NC + 1];
            if (!over) return kNoDest;
        }
        uint64_t F[KP];
#pragma unroll
        for (int j = 0; j < KP; ++j) {
            F[j] = __ballot(w0 <= s0[j]);
            if (NC > 1) F[j] &= __ballot(w1 <= s1[j]);
        }
        auto feasible = [&](int32_t q) -> bool {                     // q uniform, 0 <= q < k
            uint64_t m = F[0];
#pragma unroll
            for (int j = 1; j < KP; ++j)
                if ((q >> 6) == j) m = F[j];
            return (m >> (q & 63)) & 1ull;
        };
        // first feasible explicit candidate with index in [t0, t1)
        auto scan = [&](int t0, int t1) -> int32_t {
            const int e = t1 < 64 ? t1 : 64;
            for (int t = t0; t < e; ++t) {
                const int32_t q = __builtin_amdgcn_readlane(q0, t);
                if (feasible(q)) return q;
            }
            for (int t = t0 > 64 ? t0 : 64; t < t1; ++t) {
                const int32_t q = cand_at(x, s, off, t);
                if (feasible(q)) return q;
            }
            return kNoDest;
        };
        const bool zflag = nbz < 0;
        if (!zflag) return scan(0, nc);
        const int nbefore = nbz & 0x7fffffff;
        if (const int32_t q = scan(0, nbefore); q >= 0) return q;
        // zero group: feasible parts other than pv outside the explicit list
        uint64_t listed[KP];
#pragma unroll
        for (int j = 0; j < KP; ++j) listed[j] = (pv >> 6) == j ? 1ull << (pv & 63) : 0ull;
        for (int t = 0; t < nc; ++t) {
            const int32_t q = t < 64 ? __builtin_amdgcn_readlane(q0, t) : cand_at(x, s, off, t);
#pragma unroll
            for (int j = 0; j < KP; ++j)
                if ((q >> 6) == j) listed[j] |= 1ull << (q & 63);
        }
#pragma unroll
        for (int j = 0; j < KP; ++j)
            if (const uint64_t m = F[j] & ~listed[j]) return 64 * j + __builtin_ctzll(m);
        return scan(nbefore, nc);
    }
    RK_INL void apply(const Ctx&, int32_t, int32_t d, int32_t pv, int32_t w0, int32_t w1) {
        const int lane = __lane_id();
#pragma unroll
        for (int j = 0; j < KP; ++j) {
            if ((d >> 6) == j && lane == (d & 63)) {
                s0[j] -= w0;
                if (NC > 1) s1[j] -= w1;
            }
            if ((pv >> 6) == j && lane == (pv & 63)) {
                s0[j] += w0;
                if (NC > 1) s1[j] += w1;
            }
        }
    }
};

// Walk window slots [s0 + first, s0 + cnt) (cnt <= 64) with wave 0: applies
// each move to the running slack S and writes each slot's decision to w.dst.
template <class S>
RK_INL void resolve_chunk(const Ctx& x, S& st, int s0, int cnt, int first = 0) {
    const int lane = __lane_id();
    const int s = s0 + lane;
    int32_t v = 0, pv = 0, nc = 0, nbz = 0, off = 0, w0 = 0, w1 = 0;
    if (lane < cnt) {
        const int4 r = x.w.rec[s];
        v = r.x;
        pv = r.y;
        nc = r.z;
        nbz = r.w;
        off = x.w.off[s];
        w0 = x.w.w0[s];
        w1 = x.w.w1[s];
    }
    int32_t mine = kNoDest;
    // the staged candidates of the next mover are read one step ahead (they
    // do not depend on the slack); the pool is read at use, when nc > kCL
    int32_t qs = x.w.cl[(s0 + first) * kCL + (lane & (kCL - 1))];
    for (int i = first; i < cnt; ++i) {
        const int32_t vi = __builtin_amdgcn_readlane(v, i);
        const int32_t pvi = __builtin_amdgcn_readlane(pv, i);
        const int32_t w0i = __builtin_amdgcn_readlane(w0, i);
        const int32_t w1i = __builtin_amdgcn_readlane(w1, i);
        const int32_t nbzi = __builtin_amdgcn_readlane(nbz, i);
        const int32_t nci = __builtin_amdgcn_readlane(nc, i);
        const int32_t offi = __builtin_amdgcn_readlane(off, i);
        int32_t q0 = lane < nci ? qs : -1;
        if (nci > kCL && lane >= kCL && lane < nci) q0 = x.a->pool[offi + lane];
        if (i + 1 < cnt) qs = x.w.cl[(s0 + i + 1) * kCL + (lane & (kCL - 1))];
        const int di = st.decide(x, s0 + i, vi, pvi, nci, nbzi, offi, w0i, w1i, q0);
        if (lane == i) mine = di;
        if (di >= 0) st.apply(x, vi, di, pvi, w0i, w1i);
    }
    if (lane >= first && lane < cnt) x.w.dst[s] = mine;
}

constexpr int kSpecB = 8;    // vertex weights below 2^kSpecB in a speculative chunk
constexpr int kSpecRounds = 8;   // speculative rounds per chunk; the walk finishes the rest

// Exact speculative resolution of window slots [s0, s0+cnt) (cnt <= 64) by
// wave 0 in the cut-optimisation pass, for ncon == NC <= 2, lists of at most
// kCL candidates and no zero group.  Returns the number of leading slots it
// resolved: 0 when the chunk does not qualify, fewer than cnt when
// kSpecRounds rounds did not finish it (the walk takes over from there).
//
// Rounds start at the first unresolved lane f, whose prefix is exact.  Every
// lane >= f decides against the slack at f; the lanes that move set their bit
// in the per-part masks m_in[dest] and m_out[source].  Each lane then rebuilds
// the exact slack it would see after the lanes before it (within the round)
// applied their guesses: the weight moved in or out of a part by lanes < i is
// a popcount of that part's mask restricted to lanes < i, summed over the
// bit planes of the weights.  If every candidate the lane looked at keeps its
// verdict under that slack, its guess is the sequential decision (given that
// the lanes before it are).  The lanes before the first failing one u commit
// their moves; the next round starts at u, so a round always commits at least
// one lane and a chunk takes one round per wrong guess.  Committed moves are
// in the LDS slack, so the walk can continue from the first unresolved slot.
template <int NC>
RK_INL int spec_chunk(const Ctx& x, uint64_t* m_in, uint64_t* m_out, int s0, int cnt) {
    const int lane = __lane_id();
    const int s = s0 + lane;
    const bool act = lane < cnt;
    int32_t pv = 0, nc = 0, nbz = 0, w0 = 0, w1 = 0;
    if (act) {
        const int4 r = x.w.rec[s];
        pv = r.y;
        nc = r.z;
        nbz = r.w;
        w0 = x.w.w0[s];
        if (NC > 1) w1 = x.w.w1[s];
    }
    if (__ballot(act && (nc > kCL || nbz < 0 || (uint32_t)(w0 | w1) >= (1u << kSpecB)))) return 0;
    // bit planes of the weights: P0[b] / P1[b] = lanes whose weight has bit b
    int B = 0;
    uint64_t P0[kSpecB], P1[kSpecB];
#pragma unroll
    for (int b = 0; b < kSpecB; ++b) {
        P0[b] = __ballot((w0 >> b) & 1);
        P1[b] = NC > 1 ? __ballot((w1 >> b) & 1) : 0ull;
        if (P0[b] | P1[b]) B = b + 1;
    }
    int32_t c[kCL];
    {
        const int4 ca = *(const int4*)&x.w.cl[s * kCL];
        const int4 cb = *(const int4*)&x.w.cl[s * kCL + 4];
        const int32_t cc[kCL] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
#pragma unroll
        for (int t = 0; t < kCL; ++t) c[t] = t < nc ? cc[t] : 0;
    }
    int32_t* S = x.slack;
    int32_t d = kNoDest;
    const uint64_t bit = 1ull << lane, below = bit - 1ull;
    int f = 0;
    for (int round = 0; f < cnt && round < kSpecRounds; ++round) {
        const bool live = act && lane >= f;
        // 1. guess against the slack at f
        int32_t sv0[kCL], sv1[kCL];
#pragma unroll
        for (int t = 0; t < kCL; ++t) {
            sv0[t] = sv1[t] = 0;
            if (live && t < nc) {
                sv0[t] = S[c[t] * NC];
                if (NC > 1) sv1[t] = S[c[t] * NC + 1];
            }
        }
        int T = -1;
#pragma unroll
        for (int t = kCL - 1; t >= 0; --t)
            if (live && t < nc && w0 <= sv0[t] && (NC < 2 || w1 <= sv1[t])) T = t;
        int32_t g = kNoDest;
#pragma unroll
        for (int t = 0; t < kCL; ++t)
            if (t == T) g = c[t];
        // 2. announce the guessed moves
        if (live && g >= 0) {
            atomicOr((unsigned long long*)&m_in[g], (unsigned long long)bit);
            atomicOr((unsigned long long*)&m_out[pv], (unsigned long long)bit);
        }
        // 3. re-check every candidate up to the guess under the exact prefix
        const int last = T >= 0 ? T : nc - 1;
        bool ok = true;
#pragma unroll
        for (int t = 0; t < kCL; ++t) {
            const bool q = live && t <= last;
            if (__ballot(q) == 0) break;
            uint64_t mi = 0, mo = 0;
            if (q) {
                mi = m_in[c[t]] & below;
                mo = m_out[c[t]] & below;
            }
            long long n0 = 0, n1 = 0;                      // net weight moved in
#pragma unroll
            for (int b = 0; b < kSpecB; ++b) {             // (__popcll is unsigned: cast before subtracting)
                if (b >= B) break;
                n0 += (long long)((int)__popcll(mi & P0[b]) - (int)__popcll(mo & P0[b])) * (1ll << b);
                if (NC > 1) n1 += (long long)((int)__popcll(mi & P1[b]) - (int)__popcll(mo & P1[b])) * (1ll << b);
            }
            if (q) {
                const bool fits_t = w0 <= (long long)sv0[t] - n0 && (NC < 2 || w1 <= (long long)sv1[t] - n1);
                if (fits_t != (t == T)) ok = false;
            }
        }
        // 4. commit the lanes before the first failure, clear the masks
        const uint64_t bad = __ballot(live && !ok);
        const int u = bad ? __builtin_ctzll(bad) : cnt;
        if (live && lane < u) {
            d = g;
            if (g >= 0) {
                atomicSub(&S[g * NC], w0);
                atomicAdd(&S[pv * NC], w0);
                if (NC > 1) {
                    atomicSub(&S[g * NC + 1], w1);
                    atomicAdd(&S[pv * NC + 1], w1);
                }
            }
        }
        if (live && g >= 0) {
            m_in[g] = 0;
            m_out[pv] = 0;
        }
        f = u;
    }
    if (act && lane < f) x.w.dst[s] = d;
    return f;
}

// L: the k*ncon slack/threshold tables live in LDS (k*ncon <= kLdsKC);
// otherwise they use a.gscratch.  KP > 0: the walk keeps the slack in
// registers (k <= 64*KP, ncon == NC <= 2).  NC > 0: cut-optimisation chunks
// are resolved speculatively where spec_chunk applies.
template <bool L, int KP, int NC>
__global__ __launch_bounds__(kAT) void k_admit(AdmitArgs arg, const AdmitArgs* __restrict__ table) {
    // a batched level passes one argument block per graph (workgroup)
    AdmitArgs a = arg;
    if (table) {
        a = table[blockIdx.x];
        a.mode = arg.mode;
    }
    static_assert(KP == 0 || L, "register slack needs the LDS tables");
    using Slack = std::conditional_t<(KP > 0), RegSlack<(KP > 0 ? KP : 1), NC>, TableSlack<L>>;
    extern __shared__ int32_t s_dyn[];
    __shared__ int s_w[kAT / 64];
    if (*a.done) return;
    const int tid = threadIdx.x;
    const int kc = a.k * a.ncon;
    int32_t *slack, *fm, *om, *uni;
    uint64_t *m_in = nullptr, *m_out = nullptr;
    if constexpr (L) {
        slack = s_dyn;
        fm = s_dyn + kc;
        om = s_dyn + 2 * kc;
        uni = s_dyn + ((3 * kc + 3) & ~3);   // 16-byte aligned
        if constexpr (NC > 0) {
            m_in = (uint64_t*)uni;
            m_out = m_in + a.k;
            uni += 4 * a.k;
            for (int i = tid; i < 2 * a.k; i += kAT) m_in[i] = 0;
        }
    } else {
        slack = a.gscratch;
        fm = a.gscratch + kc;
        om = a.gscratch + 2 * kc;
        uni = s_dyn;
    }
    for (int i = tid; i < kc; i += kAT) {
        const int32_t f = clamp_thr(a.fmax[i]);
        fm[i] = f;
        om[i] = clamp_thr(a.omax[i]);
        slack[i] = f - a.pw[i];
    }
    // 1. ordered gather of the per-tile records into the dense list a.mov.
    //    The gain kernel set one bit per tile with movers: the set bits give
    //    the ordered list of those tiles (the words are cleared for the next
    //    pass); per window of kCntWin listed tiles, their counts become
    //    offsets by a block scan and every thread copies records located by a
    //    binary search over the offsets.
    const int nwords = (a.ntiles + 31) / 32;
    const int wper = (nwords + kAT - 1) / kAT;
    const int wd0 = min(tid * wper, nwords), wd1 = min(wd0 + wper, nwords);
    int nz = 0;
    for (int wd = wd0; wd < wd1; ++wd) nz += __popc(a.tile_bits[wd]);
    int nlisted;
    {
        int off = block_excl_scan_int(nz, s_w, &nlisted);
        for (int wd = wd0; wd < wd1; ++wd) {
            uint32_t bits = a.tile_bits[wd];
            if (bits) a.tile_bits[wd] = 0;
            while (bits) {
                a.tile_list[off++] = wd * 32 + __ffs(bits) - 1;
                bits &= bits - 1;
            }
        }
    }
    __threadfence();
    __syncthreads();
    int32_t* ids = uni;             // window of listed tiles
    int32_t* offs = uni + kCntWin;  // their record counts, then offsets
    int nmov = 0;
    for (int lw = 0; lw < nlisted; lw += kCntWin) {
        const int nt = min(kCntWin, nlisted - lw);
        for (int i = tid; i < nt; i += kAT) {
            const int32_t tl = a.tile_list[lw + i];
            ids[i] = tl;
            offs[i] = a.tile_cnt[tl];
        }
        __syncthreads();
        const int per = (nt + kAT - 1) / kAT;
        const int t0 = min(tid * per, nt), t1 = min(t0 + per, nt);
        int cnt = 0;
        for (int t = t0; t < t1; ++t) cnt += offs[t];
        int wtot;
        int off = block_excl_scan_int(cnt, s_w, &wtot);
        for (int t = t0; t < t1; ++t) {
            const int c = offs[t];
            offs[t] = off;
            off += c;
        }
        __syncthreads();
        // record i of the window comes from the last listed tile whose offset is <= i
        for (int i = tid; i < wtot; i += kAT) {
            int lo = 0, hi = nt - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (offs[mid] <= i) lo = mid;
                else hi = mid - 1;
            }
            a.mov[nmov + i] = a.tile_rec[(int64_t)ids[lo] * kTile + (i - offs[lo])];
        }
        nmov += wtot;
        __syncthreads();
    }
    __threadfence();
    __syncthreads();
    // 2. windows of kWin movers: stage in LDS, walk with wave 0, then all
    //    threads write the window's moves and new parts in vertex order
    Win w;
    w.rec = (int4*)uni;
    w.off = uni + 4 * kWin;
    w.w0 = w.off + kWin;
    w.w1 = w.w0 + kWin;
    w.dst = w.w1 + kWin;
    w.cl = w.dst + kWin;
    const Ctx x{&a, slack, fm, om, w};
    int nout = 0;
    for (int W0 = 0; W0 < nmov; W0 += kWin) {
        const int wn = min(kWin, nmov - W0);
        int4 r = make_int4(-1, 0, 0, 0);
        if (tid < wn) {
            r = a.mov[W0 + tid];
            const int32_t off = a.xadj[r.x];
            w.rec[tid] = r;
            w.off[tid] = off;
            w.w0[tid] = a.vwgt ? a.vwgt[(int64_t)r.x * a.ncon] : 1;
            w.w1[tid] = a.ncon > 1 ? (a.vwgt ? a.vwgt[(int64_t)r.x * a.ncon + 1] : 1) : 0;
            const int lim = r.z < kCL ? r.z : kCL;
            for (int t = 0; t < lim; ++t) w.cl[tid * kCL + t] = a.pool[off + t];
        }
        __syncthreads();
        if ((tid >> 6) == 0)
            for (int s0 = 0; s0 < wn; s0 += 64) {
                const int cnt = min(64, wn - s0);
                int first = 0;
                if constexpr (NC > 0)
                    if (a.mode == 1) first = spec_chunk<NC>(x, m_in, m_out, s0, cnt);
                if (first == cnt) continue;
                Slack st;
                st.load(x);
                resolve_chunk(x, st, s0, cnt, first);
                st.store(x);
            }
        __syncthreads();
        const int d = tid < wn ? w.dst[tid] : kNoDest;
        int tot;
        const int pos = block_flag_scan(d >= 0, s_w, &tot);
        if (d >= 0) {
            a.moves[nout + pos] = make_int2(r.x, d);
            a.part[r.x] = d;
        }
        nout += tot;
        __syncthreads();
    }
    if constexpr (!L) __threadfence();
    __syncthreads();
    for (int i = tid; i < kc; i += kAT) a.pw[i] = fm[i] - slack[i];
    if (tid == 0) {
        *a.nmoves = nout;
        *a.passes += 1;
        if (a.mode == 1 && nout == 0) *a.done = 1;
    }
}

}  // namespace

namespace {

// Launches k_admit for graphs of k*ncon = kc at most: one workgroup with the
// arguments aa, or one per graph of a batched level with arguments from table.
void launch_admit(const AdmitArgs& aa, const AdmitArgs* table, int grid, int32_t k, int32_t ncon, hipStream_t s) {
    const int kc = k * ncon;
    const size_t win = (size_t)kWinInts * sizeof(int32_t);
    const size_t cntb = 2 * (size_t)kCntWin * sizeof(int32_t);
    const size_t uni = win > cntb ? win : cntb;
    if (kc <= kLdsKC) {
        const size_t base = (size_t)((3 * kc + 3) & ~3) * sizeof(int32_t) + uni;
        // the speculative path's part masks ride along while LDS stays within 64 KiB
        const int nc2 = ncon <= 2 && base + 16 * (size_t)k <= kLdsBytes ? ncon : 0;
        const size_t lds = base + (nc2 ? 16 * (size_t)k : 0);
        const int kp = nc2 == 0 ? 0 : (k <= 64 ? 1 : (k <= 128 ? 2 : (k <= 256 ? 4 : 0)));
        switch (kp * 4 + nc2) {
        case 4 + 1: hipLaunchKernelGGL((k_admit<true, 1, 1>), dim3(grid), dim3(kAT), lds, s, aa, table); break;
        case 4 + 2: hipLaunchKernelGGL((k_admit<true, 1, 2>), dim3(grid), dim3(kAT), lds, s, aa, table); break;
        case 8 + 1: hipLaunchKernelGGL((k_admit<true, 2, 1>), dim3(grid), dim3(kAT), lds, s, aa, table); break;
        case 8 + 2: hipLaunchKernelGGL((k_admit<true, 2, 2>), dim3(grid), dim3(kAT), lds, s, aa, table); break;
        case 16 + 1: hipLaunchKernelGGL((k_admit<true, 4, 1>), dim3(grid), dim3(kAT), lds, s, aa, table); break;
        case 16 + 2: hipLaunchKernelGGL((k_admit<true, 4, 2>), dim3(grid), dim3(kAT), lds, s, aa, table); break;
        case 1: hipLaunchKernelGGL((k_admit<true, 0, 1>), dim3(grid), dim3(kAT), lds, s, aa, table); break;
        case 2: hipLaunchKernelGGL((k_admit<true, 0, 2>), dim3(grid), dim3(kAT), lds, s, aa, table); break;
        default: hipLaunchKernelGGL((k_admit<true, 0, 0>), dim3(grid), dim3(kAT), lds, s, aa, table); break;
        }
    } else {
        hipLaunchKernelGGL((k_admit<false, 0, 0>), dim3(grid), dim3(kAT), uni, s, aa, table);
    }
    RK_HIP(hipGetLastError());
}

}  // namespace

void admit_pass(const DevGraph& g, int32_t k, int mode, int ntiles, RefineBufs& b, int32_t* part, hipStream_t s) {
    const AdmitArgs aa{mode, k, g.ncon, ntiles, b.tile_cnt, b.tile_bits, b.tile_off, b.tile_rec, b.mov, b.pool,
                       g.xadj, g.vwgt, b.fmax, b.omax, b.pw, part, b.moves, b.nmoves, b.done, b.passes,
                       b.adm_scratch};
    launch_admit(aa, nullptr, 1, k, g.ncon, s);
}

void admit_pass_batch(int mode, int ngraphs, const AdmitArgs* table, int32_t kmax, int32_t ncon, hipStream_t s) {
    if (ngraphs <= 0) return;
    AdmitArgs aa{};
    aa.mode = mode;
    launch_admit(aa, table, ngraphs, kmax, ncon, s);
}

}  // namespace rk
