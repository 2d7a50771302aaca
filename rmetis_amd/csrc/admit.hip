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
//  * all 256 threads list the tiles with movers (a bitmap kept by the gain
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
#include "tuning.hpp"

namespace rk {

namespace {

constexpr int kAT = 256;        // threads per admission workgroup
constexpr int kLdsKC = 2048;    // k*ncon entries of slack/thresholds kept in LDS
constexpr int kCL = 32;         // candidates per mover staged in LDS
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
    int mode;             // 0 balance repair, 1 cut optimisation
    int32_t* slack;       // fmax - pw, running
    const int32_t* fm;    // clamped fmax
    const int32_t* om;    // clamped omax
    Win w;
    int32_t* wv;          // [128] LDS: speculating lanes' weights (constraint 0, then 1)
    uint64_t* pl;         // [64] LDS: their bit planes (constraint 0, then 1)
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
    if (x.mode == 0) {                                               // greedy.rs:69-74
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
    // zero group: the first feasible part other than pv outside the explicit
    // list; the candidates ahead of it were all infeasible above, so only
    // those after it can hide a feasible listed part
    auto listed_after = [&](int32_t p) __attribute__((always_inline)) -> bool {
        for (int tb = (nbefore / 64) * 64; tb < nc; tb += 64) {
            const int t = tb + lane;
            const int32_t q = t < nc ? (tb == 0 ? q0 : cand_at(x, s, off, t)) : -1;
            if (__ballot(t >= nbefore && q == p)) return true;
        }
        return false;
    };
    for (int qb = 0; qb < a.k; qb += 64) {
        const int32_t q = qb + lane;
        for (uint64_t m = __ballot(q < a.k && q != pv && fits(x, q, v, w0, w1)); m; m &= m - 1) {
            const int32_t p = qb + __builtin_ctzll(m);
            if (!listed_after(p)) return p;
        }
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

// Up to four values of one kind held in named registers.  Element access
// by index is a chain of selects on values (an array indexed at run time
// would be placed in scratch memory).
template <typename T, int N = 4>
struct Quad {
    T e0{}, e1{}, e2{}, e3{};
    RK_INL T get(int j) const {
        if constexpr (N == 1) return e0;
        else if constexpr (N == 2) return j == 0 ? e0 : e1;
        else return j == 0 ? e0 : (j == 1 ? e1 : (j == 2 ? e2 : e3));
    }
    RK_INL void set(int j, T v) {
        e0 = j == 0 ? v : e0;
        e1 = j == 1 ? v : e1;
        e2 = j == 2 ? v : e2;
        e3 = j == 3 ? v : e3;
    }
};

// Running slack in wave 0's registers for k <= 64*KP and ncon == NC <= 2:
// lane l of block j of s0 / s1 holds part 64*j + l (INT32_MIN past k, never
// feasible).  A mover's feasible parts are one compare per register block
// and constraint; no memory access sits on the mover-to-mover chain.
template <int KP, int NC>
struct RegSlack {
    static_assert(KP >= 1 && KP <= 4, "at most four register blocks");
    Quad<int32_t, KP> s0, s1;

    RK_INL void load(const Ctx& x) {
        const int lane = __lane_id(), k = x.a->k;
#pragma unroll
        for (int j = 0; j < KP; ++j) {
            const int p = 64 * j + lane;
            s0.set(j, p < k ? x.slack[p * NC] : INT32_MIN);
            s1.set(j, p < k && NC > 1 ? x.slack[p * NC + 1] : INT32_MIN);
        }
    }
    RK_INL void store(const Ctx& x) const {
        const int lane = __lane_id(), k = x.a->k;
#pragma unroll
        for (int j = 0; j < KP; ++j) {
            const int p = 64 * j + lane;
            if (p < k) {
                x.slack[p * NC] = s0.get(j);
                if (NC > 1) x.slack[p * NC + 1] = s1.get(j);
            }
        }
    }
    // uniform part p: its slack in constraint c
    RK_INL int32_t at(int32_t p, int c) const {
        const int32_t r = c == 0 ? s0.get(p >> 6) : s1.get(p >> 6);
        return __builtin_amdgcn_readlane(r, p & 63);
    }
    // F holds one feasibility bit per part; the candidates are tested 64 at a
    // time (lane t: candidate t, the first 64 already in q0, the rest loaded
    // from the pool in coalesced blocks) and a ballot picks the first one in
    // admission order.
    RK_INL int decide(const Ctx& x, int s, int32_t, int32_t pv, int32_t nc, int32_t nbz, int32_t off, int32_t w0,
                      int32_t w1, int32_t q0) const {
        const AdmitArgs& a = *x.a;
        const int lane = __lane_id();
        if (x.mode == 0) {                                           // greedy.rs:69-74
            bool over = (long long)x.fm[pv * NC] - at(pv, 0) > (long long)x.om[pv * NC];
            if (NC > 1) over = over || (long long)x.fm[pv * NC + 1] - at(pv, 1) > (long long)x.om[pv * NC + 1];
            if (!over) return kNoDest;
        }
        Quad<uint64_t, KP> F;
#pragma unroll
        for (int j = 0; j < KP; ++j) {
            uint64_t m = __ballot(w0 <= s0.get(j));
            if (NC > 1) m &= __ballot(w1 <= s1.get(j));
            F.set(j, m);
        }
        auto feasible = [&](int32_t q) __attribute__((always_inline)) -> bool {   // per lane; q < 0: false
            return q >= 0 && ((F.get(q >> 6) >> (q & 63)) & 1ull);
        };
        auto cand = [&](int tb) __attribute__((always_inline)) -> int32_t {       // lane's candidate tb + lane, or -1
            const int t = tb + lane;
            if (t >= nc) return -1;
            return tb == 0 ? q0 : a.pool[off + t];
        };
        // first feasible explicit candidate with index in [t0, t1): a few
        // candidates one at a time in scalar code, longer ranges 64 at a time
        auto scan = [&](int t0, int t1) __attribute__((always_inline)) -> int32_t {
            if (t1 <= 64 && t1 - t0 <= 4) {
                for (int t = t0; t < t1; ++t) {
                    const int32_t q = __builtin_amdgcn_readlane(q0, t);
                    if ((F.get(q >> 6) >> (q & 63)) & 1ull) return q;
                }
                return kNoDest;
            }
            for (int tb = (t0 / 64) * 64; tb < t1; tb += 64) {
                const int t = tb + lane;
                const int32_t q = t >= t0 && t < t1 ? cand(tb) : -1;
                if (const uint64_t m = __ballot(feasible(q))) return __builtin_amdgcn_readlane(q, __builtin_ctzll(m));
            }
            return kNoDest;
        };
        const bool zflag = nbz < 0;
        if (!zflag) return scan(0, nc);
        const int nbefore = nbz & 0x7fffffff;
        if (const int32_t q = scan(0, nbefore); q >= 0) return q;
        // zero group: the first feasible part other than pv outside the
        // explicit list.  The candidates ahead of the zero group were all
        // infeasible just above, so only those after it can hide a feasible
        // listed part.
        auto listed_after = [&](int32_t p) __attribute__((always_inline)) -> bool {
            for (int tb = (nbefore / 64) * 64; tb < nc; tb += 64) {
                const int t = tb + lane;
                if (__ballot(t >= nbefore && cand(tb) == p)) return true;
            }
            return false;
        };
        for (int j = 0; j < KP; ++j) {
            uint64_t m = F.get(j);
            if ((pv >> 6) == j) m &= ~(1ull << (pv & 63));
            for (; m; m &= m - 1) {
                const int32_t p = 64 * j + __builtin_ctzll(m);
                if (!listed_after(p)) return p;
            }
        }
        return scan(nbefore, nc);
    }
    RK_INL void apply(const Ctx&, int32_t, int32_t d, int32_t pv, int32_t w0, int32_t w1) {
        const int lane = __lane_id();
#pragma unroll
        for (int j = 0; j < KP; ++j) {
            int32_t a0 = s0.get(j), a1 = s1.get(j);
            if ((d >> 6) == j && lane == (d & 63)) {
                a0 -= w0;
                a1 -= NC > 1 ? w1 : 0;
            }
            if ((pv >> 6) == j && lane == (pv & 63)) {
                a0 += w0;
                a1 += NC > 1 ? w1 : 0;
            }
            s0.set(j, a0);
            s1.set(j, a1);
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
    // the first 64 candidates of the next mover are read one step ahead (they
    // do not depend on the slack): the staged ones from LDS into qs, the rest
    // from the pool into qg.  The two stay separate loads, selected by value
    // at use (a select of the two addresses would make one flat load).
    int32_t qs = x.w.cl[(s0 + first) * kCL + (lane & (kCL - 1))], qg = -1;
    {
        const int32_t nc0 = __builtin_amdgcn_readlane(nc, first);
        if (nc0 > kCL && lane >= kCL && lane < nc0) qg = x.a->pool[__builtin_amdgcn_readlane(off, first) + lane];
    }
    for (int i = first; i < cnt; ++i) {
        const int32_t vi = __builtin_amdgcn_readlane(v, i);
        const int32_t pvi = __builtin_amdgcn_readlane(pv, i);
        const int32_t w0i = __builtin_amdgcn_readlane(w0, i);
        const int32_t w1i = __builtin_amdgcn_readlane(w1, i);
        const int32_t nbzi = __builtin_amdgcn_readlane(nbz, i);
        const int32_t nci = __builtin_amdgcn_readlane(nc, i);
        const int32_t offi = __builtin_amdgcn_readlane(off, i);
        const int32_t q0 = lane < nci ? (lane < kCL ? qs : qg) : -1;
        if (i + 1 < cnt) {
            const int32_t ncn = __builtin_amdgcn_readlane(nc, i + 1);
            qs = x.w.cl[(s0 + i + 1) * kCL + (lane & (kCL - 1))];
            qg = -1;
            if (ncn > kCL && lane >= kCL && lane < ncn) qg = x.a->pool[__builtin_amdgcn_readlane(off, i + 1) + lane];
        }
        const int di = st.decide(x, s0 + i, vi, pvi, nci, nbzi, offi, w0i, w1i, q0);
        if (lane == i) mine = di;
        if (di >= 0) st.apply(x, vi, di, pvi, w0i, w1i);
    }
    if (lane >= first && lane < cnt) x.w.dst[s] = mine;
}

constexpr int kSpecRounds = 8;   // speculative rounds per chunk; the walk finishes the rest
constexpr int kSpecMinRun = 4;   // shortest run of slots between barriers worth a speculative call

// Net weight moved into a part by the lanes of masks mi (in) and mo (out):
// one lane at a time when they are few, else by the bit planes of the lanes'
// weights (weighted popcounts; __popcll is unsigned, so cast first).
template <int NC>
RK_INL void net_inflow(const Ctx& x, uint64_t mi, uint64_t mo, int B, long long& n0, long long& n1) {
    n0 = 0;
    n1 = 0;
    if (B == 1) {                                              // every weight is 1
        n0 = (long long)((int)__popcll(mi) - (int)__popcll(mo));
        n1 = n0;
    } else if ((int)(__popcll(mi) + __popcll(mo)) <= B) {
        for (uint64_t m = mi; m; m &= m - 1) {
            const int j = __builtin_ctzll(m);
            n0 += x.wv[j];
            if (NC > 1) n1 += x.wv[64 + j];
        }
        for (uint64_t m = mo; m; m &= m - 1) {
            const int j = __builtin_ctzll(m);
            n0 -= x.wv[j];
            if (NC > 1) n1 -= x.wv[64 + j];
        }
    } else {
        for (int b = 0; b < B; ++b) {
            const uint64_t p0 = x.pl[b];
            n0 += (long long)((int)__popcll(mi & p0) - (int)__popcll(mo & p0)) * (1ll << b);
            if (NC > 1) {
                const uint64_t p1 = x.pl[32 + b];
                n1 += (long long)((int)__popcll(mi & p1) - (int)__popcll(mo & p1)) * (1ll << b);
            }
        }
    }
}

// Speculative rounds over slots [from, lim) of a chunk whose lists hold at
// most CL candidates (CL <= kCL, a multiple of 4); writes the decisions of
// the resolved slots and returns the first unresolved one.
template <int NC, int CL>
RK_INL int spec_rounds(const Ctx& x, uint64_t* m_in, uint64_t* m_out, int s0, bool act, int32_t pv, int32_t nc,
                       int32_t w0, int32_t w1, int from, int lim, int B) {
    const int lane = __lane_id();
    const int s = s0 + lane;
    // the staged candidate list, one register per position
    int32_t c[CL];
    {
        const int4* cl4 = (const int4*)&x.w.cl[(act ? s : s0) * kCL];
#pragma unroll
        for (int t4 = 0; t4 < CL / 4; ++t4) {
            const int4 v = 4 * t4 < nc ? cl4[t4] : make_int4(0, 0, 0, 0);
            c[4 * t4] = 4 * t4 < nc ? v.x : 0;            // entries past the list are not staged
            c[4 * t4 + 1] = 4 * t4 + 1 < nc ? v.y : 0;
            c[4 * t4 + 2] = 4 * t4 + 2 < nc ? v.z : 0;
            c[4 * t4 + 3] = 4 * t4 + 3 < nc ? v.w : 0;
        }
    }
    int32_t* S = x.slack;
    auto fits_at = [&](int32_t q, long long n0, long long n1) __attribute__((always_inline)) -> bool {   // q's slack less the net inflow
        return w0 <= (long long)S[q * NC] - n0 && (NC < 2 || w1 <= (long long)S[q * NC + 1] - n1);
    };
    int32_t d = kNoDest;
    const uint64_t bit = 1ull << lane, below = bit - 1ull;
    int f = from;
    for (int round = 0; f < lim && round < kSpecRounds; ++round) {
        const bool live = act && lane >= f;
        // 1. guess against the slack at f: the first candidate that fits
        int T = -1;
#pragma unroll
        for (int t = CL - 1; t >= 0; --t)
            if (live && t < nc && fits_at(c[t], 0, 0)) T = t;
        int32_t g = kNoDest;
#pragma unroll
        for (int t = 0; t < CL; ++t)
            if (t == T) g = c[t];
        // 2. announce the guessed moves
        if (live && g >= 0) {
            atomicOr((unsigned long long*)&m_in[g], (unsigned long long)bit);
            atomicOr((unsigned long long*)&m_out[pv], (unsigned long long)bit);
        }
        // 3. re-check, under the exact prefix, every candidate up to the guess
        //    that a lane before this one moved into or out of (the others keep
        //    the verdict they had at f): first mark them, then check them one
        //    at a time (usually none or one per lane)
        const int last = T >= 0 ? T : nc - 1;
        uint32_t touched = 0;
#pragma unroll
        for (int t = 0; t < CL; ++t)
            if (live && t <= last && ((m_in[c[t]] | m_out[c[t]]) & below)) touched |= 1u << t;
        bool ok = true;
        for (; touched; touched &= touched - 1) {
            const int t = __builtin_ctz(touched);
            const int32_t q = x.w.cl[s * kCL + t];
            long long n0, n1;
            net_inflow<NC>(x, m_in[q] & below, m_out[q] & below, B, n0, n1);
            if (fits_at(q, n0, n1) != (t == T)) ok = false;
        }
        // 4. commit the lanes before the first failure, clear the masks
        const uint64_t bad = __ballot(live && !ok);
        const int u = bad ? __builtin_ctzll(bad) : lim;
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

// Exact speculative resolution of window slots [s0+from, s0+cnt) (cnt <= 64)
// by wave 0 in the cut-optimisation pass, for ncon == NC <= 2.  A slot whose
// mover has more than kCL candidates or a zero group is a barrier: the slots before the first barrier are speculated, and
// the function returns the first slot f it did not resolve.  *walk_end is
// where the walk that takes over from f should stop: after the run of
// barriers at f and any short gaps between barriers (fewer than
// kSpecMinRun slots, not worth a speculative call), or, as -cnt, at cnt
// when the rounds ran out before a barrier.
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
RK_INL int spec_chunk(const Ctx& x, uint64_t* m_in, uint64_t* m_out, int s0, int cnt, int from, int* walk_end) {
    const int lane = __lane_id();
    const int s = s0 + lane;
    int32_t pv = 0, nc = 0, nbz = 0, w0 = 0, w1 = 0;
    if (lane < cnt) {
        const int4 r = x.w.rec[s];
        pv = r.y;
        nc = r.z;
        nbz = r.w;
        w0 = x.w.w0[s];
        if (NC > 1) w1 = x.w.w1[s];
    }
    const uint64_t dq =
        __ballot(lane >= from && lane < cnt && (nc > kCL || nbz < 0));
    const int lim = dq ? __builtin_ctzll(dq) : cnt;
    {
        int e = lim;   // barriers and short gaps from lim on
        while (e < cnt) {
            if ((dq >> e) & 1ull) {
                ++e;
                continue;
            }
            const uint64_t rest = dq >> e;
            const int run = rest ? __builtin_ctzll(rest) : cnt - e;
            if (run >= kSpecMinRun) break;
            e += run;
        }
        *walk_end = e;
    }
    if (lim == from) return from;
    const bool act = lane >= from && lane < lim;
    if (!act) nc = 0;
    // every lane's weights, and their bit planes pl[b] (pl[32 + b] for the
    // second constraint) = the lanes whose weight has bit b, for the net
    // inflow sums of the verification; B = planes in use
    x.wv[lane] = w0;
    if (NC > 1) x.wv[64 + lane] = w1;
    int wor = act ? (w0 | w1) : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) wor |= __shfl_xor(wor, o, 64);
    const int B = 32 - __builtin_clz((uint32_t)__builtin_amdgcn_readfirstlane(wor) | 1u);
    for (int b = 0; b < B; ++b) {
        const uint64_t m0 = __ballot(act && ((w0 >> b) & 1));
        const uint64_t m1 = NC > 1 ? __ballot(act && ((w1 >> b) & 1)) : 0ull;
        if (lane == 0) {
            x.pl[b] = m0;
            if (NC > 1) x.pl[32 + b] = m1;
        }
    }
    const int f = __ballot(act && nc > 8)
                      ? spec_rounds<NC, kCL>(x, m_in, m_out, s0, act, pv, nc, w0, w1, from, lim, B)
                      : spec_rounds<NC, 8>(x, m_in, m_out, s0, act, pv, nc, w0, w1, from, lim, B);
    if (f < lim) *walk_end = -cnt;   // rounds ran out: the walk takes the rest
    return f;
}

// L: the k*ncon slack/threshold tables live in LDS (k*ncon <= kLdsKC);
// otherwise they use a.gscratch.  KP > 0: the walk keeps the slack in
// registers (k <= 64*KP, ncon == NC <= 2).  NC > 0: cut-optimisation chunks
// are resolved speculatively where spec_chunk applies.
template <bool L, int KP, int NC>
RK_INL void admit_body(const AdmitArgs& a, const int mode, const int32_t spec_min) {
    static_assert(KP == 0 || L, "register slack needs the LDS tables");
    using Slack = std::conditional_t<(KP > 0), RegSlack<(KP > 0 ? KP : 1), NC>, TableSlack<L>>;
    extern __shared__ int32_t s_dyn[];
    __shared__ int s_w[kAT / 64];
    __shared__ int32_t s_wv[128];
    __shared__ uint64_t s_pl[64];
    if (*a.done) return;
    const int tid = threadIdx.x;
    const int kc = a.k * a.ncon;
    // tiles the next cut pass recomputes: marked here for the graphs of a
    // batched level (a.tile_dirty given), by k_mark_moves for one graph
    const bool mark = mode == 1 && a.tile_dirty != nullptr;
    if (mode == 1 && a.full && tid == 0) *a.full = 0;
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
    // the bits of tiles past ntiles (a finer level's) are not this graph's
    auto word = [&](int wd) __attribute__((always_inline)) -> uint32_t {
        const uint32_t b = a.tile_bits[wd];
        const int past = (wd + 1) * 32 - a.ntiles;
        return past > 0 ? b & (0xffffffffu >> past) : b;
    };
    int nz = 0;
    for (int wd = wd0; wd < wd1; ++wd) nz += __popc(word(wd));
    int nlisted;
    {
        int off = block_excl_scan_int(nz, s_w, &nlisted);
        for (int wd = wd0; wd < wd1; ++wd) {
            uint32_t bits = word(wd);
            while (bits) {
                a.tile_list[off++] = wd * 32 + __ffs(bits) - 1;
                bits &= bits - 1;
            }
        }
    }
    // tile_list, mov and gscratch are read back only by this workgroup: the
    // barrier's workgroup-scope release/acquire orders them (an agent-scope
    // fence would write back and invalidate the whole L2 of the XCD)
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
    const Ctx x{&a, mode, slack, fm, om, w, s_wv, s_pl};
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
            // the first kCL candidates: every load issued before the first
            // store, then written in 16-byte blocks
            const int lim = r.z < kCL ? r.z : kCL;
            const int32_t* src = a.pool + off;
            int32_t cv[kCL];
#pragma unroll
            for (int t = 0; t < kCL; ++t) cv[t] = t < lim ? src[t] : 0;
            int4* dst4 = (int4*)&w.cl[tid * kCL];
#pragma unroll
            for (int t4 = 0; t4 < kCL / 4; ++t4)
                if (4 * t4 < lim) dst4[t4] = make_int4(cv[4 * t4], cv[4 * t4 + 1], cv[4 * t4 + 2], cv[4 * t4 + 3]);
        }
        __syncthreads();
        // a speculative call that runs out of rounds (many moves conflict)
        // turns speculation off for the rest of the window.  Levels with
        // fewer than spec_min vertices per part (the tuning table; the coarsest, whose
        // vertices weigh a large share of a part's slack) conflict so often
        // that the walk alone is faster there; both are exact.
        bool spec_on = mode == 1 && (int64_t)a.ntiles * 64 >= (int64_t)spec_min * a.k;
        if ((tid >> 6) == 0)
            for (int s0 = 0; s0 < wn; s0 += 64) {
                const int cnt = min(64, wn - s0);
                int i = 0;
                while (i < cnt) {
                    int e = cnt;   // the walk takes [i, e)
                    if constexpr (NC > 0)
                        if (spec_on) {
                            i = spec_chunk<NC>(x, m_in, m_out, s0, cnt, i, &e);
                            if (i == cnt) break;
                            if (e < 0) {
                                e = -e;
                                spec_on = false;
                            }
                        }
                    Slack st;
                    st.load(x);
                    resolve_chunk(x, st, s0, e, i);
                    st.store(x);
                    i = e;
                }
            }
        __syncthreads();
        const int d = tid < wn ? w.dst[tid] : kNoDest;
        int tot;
        const int pos = block_flag_scan(d >= 0, s_w, &tot);
        if (d >= 0) {
            a.moves[nout + pos] = make_int2(r.x, d);
            a.part[r.x] = d;
            if (mark) {
                // the mover's tile and its neighbours' tiles: their gain
                // tables change; a hub sends the next pass over every tile
                auto dirty = [&](int32_t u) __attribute__((always_inline)) {
                    const int32_t t = (u - a.v0) >> 6;
                    atomicOr(&a.tile_dirty[t >> 5], 1u << (t & 31));
                };
                dirty(r.x);
                const int32_t b0 = a.xadj[r.x], b1 = a.xadj[r.x + 1];
                if (b1 - b0 > kMarkMaxDeg) *a.full = 1;
                else
                    for (int32_t j = b0; j < b1; ++j) dirty(a.adjncy[j]);
            }
        }
        nout += tot;
        __syncthreads();
    }
    __syncthreads();
    for (int i = tid; i < kc; i += kAT) a.pw[i] = fm[i] - slack[i];
    if (tid == 0) {
        *a.nmoves = nout;
        *a.passes += 1;
        if (mode == 1 && nout == 0) *a.done = 1;
    }
}

// One pass over the graph described by arg, or, for a batched level, one per
// workgroup over the graph described by table[blockIdx.x] (mode from arg).
template <bool L, int KP, int NC>
__global__ __launch_bounds__(kAT) void k_admit(AdmitArgs arg, const AdmitArgs* __restrict__ table) {
    AdmitArgs a = arg;
    if (table) a = table[blockIdx.x];
    admit_body<L, KP, NC>(a, arg.mode, arg.spec_min);
}

}  // namespace

namespace {

// Launches k_admit for graphs of k*ncon = kc at most: one workgroup with the
// arguments aa, or one per graph of a batched level with arguments from table.
void launch_admit(const AdmitArgs& aa0, const AdmitArgs* table, int grid, int32_t k, int32_t ncon, hipStream_t s) {
    AdmitArgs aa = aa0;
    aa.spec_min = tune(kTuneSpecMin);
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
                       b.adm_scratch, g.adjncy, nullptr, b.full, 0, 0};
    launch_admit(aa, nullptr, 1, k, g.ncon, s);
}

void admit_pass_batch(int mode, int ngraphs, const AdmitArgs* table, int32_t kmax, int32_t ncon, hipStream_t s) {
    if (ngraphs <= 0) return;
    AdmitArgs aa{};
    aa.mode = mode;
    launch_admit(aa, table, ngraphs, kmax, ncon, s);
}

}  // namespace rk
