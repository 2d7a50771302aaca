// admit.hip -- ordered admission of one greedy pass (refine/greedy.rs:64-123,
// 131-181) in a single workgroup.
//
// The reference visits the candidate movers in ascending vertex order; each
// one takes the first part of its gain-ordered candidate list that is
// feasible under the *running* part weights (and, in the repair pass, only if
// its own part is still overweight), then updates those weights.  Passes
// carry tens to a few thousand candidates, so this step is latency-bound:
//
//  * all 256 threads read the per-tile candidate counts, gather the records
//    into one dense ordered list, and stage windows of 256 movers (record,
//    CSR offset, vertex weights, first candidates) plus pw and the integer
//    thresholds in LDS;
//  * wave 0 resolves each window in chunks of 64 (one mover per lane).  In
//    the cut-optimisation pass every lane first decides against the
//    chunk-start weights; the chunk's tentative in/out flows A, R bound every
//    prefix weight by [pw - R, pw + A], and a lane whose decision is the same
//    at both corners (feasibility is monotone) is certified.  If all lanes
//    certify the guesses are the sequential result; otherwise, from the first
//    uncertified lane on, the chunk is solved as a fixed point: each lane
//    re-decides against the chunk-start weights plus the exact flows of the
//    lanes before it, until no lane changes (lane i is final after i+1
//    rounds, and a fixed point is the sequential result).  The repair pass,
//    long candidate lists and more than two constraints use an exact
//    one-mover-at-a-time walk in which the lanes test candidate parts in
//    parallel.
//
// Thresholds floor(tpw*ub) are stored clamped to int32: with every
// per-constraint total below 2^31 (checked by the callers) each running
// weight and weight+vw lies in [0, 2^31), so the clamped compares equal the
// reference's f64 compares (greedy.rs:70,87,144).
#include <climits>

#include "common.hpp"
#include "kernels.hpp"

namespace rk {

namespace {

constexpr int kAT = 256;        // threads per admission workgroup
constexpr int kLdsKC = 2048;    // k*ncon entries of pw/thresholds/flows kept in LDS
constexpr int kCK = 4;          // candidates per mover in the fixed-point path
constexpr int kCL = 8;          // candidates per mover staged in LDS
constexpr int kWin = 256;       // movers staged per window
constexpr int kCntWin = 4096;   // tile counts read per window
constexpr int kGateOff = -2;    // repair: source part no longer overweight
constexpr int kNoDest = -1;

struct AdmArgs {
    int32_t mode, k, ncon, ntiles;
    const int32_t* tile_cnt;
    const int4* tile_rec;
    int4* mov;
    const int32_t* pool;
    const int32_t* xadj;
    const int32_t* vwgt;
    const long long* fmax;
    const long long* omax;
    int32_t* pw;
    int32_t* part;
    int2* moves;
    int32_t* nmoves;
    int32_t* done;
    int32_t* passes;
    int32_t* gscratch;   // 4*k*ncon ints when k*ncon > kLdsKC
};

// LDS layout of the mover window (after the 5*kc weight/threshold/flow ints).
struct Win {
    int4* rec;       // [kWin]
    int32_t* off;    // [kWin]  xadj[v]
    int32_t* w0;     // [kWin]
    int32_t* w1;     // [kWin]
    int32_t* cl;     // [kWin * kCL]
};

__device__ inline int32_t clamp32(long long x) {
    return x > INT32_MAX ? INT32_MAX : (x < INT32_MIN ? INT32_MIN : (int32_t)x);
}

__device__ inline int block_excl_scan_int(int x, int* s_w, int* total) {
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

struct Ctx {
    const AdmArgs* a;
    int32_t* pw;
    const int32_t* fm;
    const int32_t* om;
    int32_t* A;
    int32_t* R;
    Win w;
};

__device__ inline int32_t vw_of(const AdmArgs& a, int32_t v, int c) {
    return a.vwgt ? a.vwgt[(int64_t)v * a.ncon + c] : 1;
}

// Candidate t of the mover in window slot s.
__device__ inline int32_t cand_at(const Ctx& x, int s, int32_t off, int t) {
    return t < kCL ? x.w.cl[s * kCL + t] : x.a->pool[off + t];
}

// Exact decision of one mover against the running weights; the 64 lanes test
// candidate parts in parallel and the first feasible one in admission order
// wins.  All arguments are wave-uniform.
__device__ int decide_seq(const Ctx& x, int s, int32_t v, int32_t pv, int32_t nc, int32_t nbz, int32_t off) {
    const AdmArgs& a = *x.a;
    const int lane = __lane_id();
    const int ncon = a.ncon;
    if (a.mode == 0) {
        bool over = false;                                           // greedy.rs:69-74
        for (int c = 0; c < ncon; ++c)
            if (x.pw[pv * ncon + c] > x.om[pv * ncon + c]) over = true;
        if (!over) return kGateOff;
    }
    auto feasible = [&](int32_t q) {                                 // greedy.rs:85-88 / 142-145
        for (int c = 0; c < ncon; ++c)
            if ((long long)x.pw[q * ncon + c] + vw_of(a, v, c) > (long long)x.fm[q * ncon + c]) return false;
        return true;
    };
    const int nbefore = nbz & 0x7fffffff;
    const bool zflag = ((uint32_t)nbz >> 31) != 0;
    for (int tb = 0; tb < nbefore; tb += 64) {
        const int t = tb + lane;
        const int32_t q = t < nbefore ? cand_at(x, s, off, t) : -1;
        const uint64_t bm = __ballot(q >= 0 && feasible(q));
        if (bm) return __shfl(q, __ffsll((unsigned long long)bm) - 1, 64);
    }
    if (zflag) {
        for (int qb = 0; qb < a.k; qb += 64) {
            const int32_t q = qb + lane;
            bool ok = q < a.k && q != pv;
            if (ok)
                for (int t = 0; t < nc; ++t)
                    if (cand_at(x, s, off, t) == q) {
                        ok = false;
                        break;
                    }
            const uint64_t bm = __ballot(ok && feasible(q));
            if (bm) return qb + __ffsll((unsigned long long)bm) - 1;
        }
    }
    for (int tb = nbefore; tb < nc; tb += 64) {
        const int t = tb + lane;
        const int32_t q = t < nc ? cand_at(x, s, off, t) : -1;
        const uint64_t bm = __ballot(q >= 0 && feasible(q));
        if (bm) return __shfl(q, __ffsll((unsigned long long)bm) - 1, 64);
    }
    return kNoDest;
}

// Resolve window slots [s0, s0+cnt) (cnt <= 64) with wave 0; returns #moves.
__device__ int resolve_chunk(const Ctx& x, int s0, int cnt, int nout) {
    const AdmArgs& a = *x.a;
    const int lane = __lane_id();
    const int ncon = a.ncon;
    const int s = s0 + lane;
    const bool act = lane < cnt;
    int32_t v = 0, pv = 0, nc = 0, nbz = 0, off = 0, w0 = 0, w1 = 0;
    if (act) {
        const int4 r = x.w.rec[s];
        v = r.x;
        pv = r.y;
        nc = r.z;
        nbz = r.w;
        off = x.w.off[s];
        w0 = x.w.w0[s];
        w1 = x.w.w1[s];
    }
    const bool simple = !act || (nc <= kCK && ((uint32_t)nbz >> 31) == 0);
    const bool fast = a.mode == 1 && ncon <= 2 && __ballot(!simple) == 0;
    if (!fast) {
        int moved = 0;
        for (int i = 0; i < cnt; ++i) {
            const int32_t vi = __builtin_amdgcn_readlane(v, i);
            const int32_t pvi = __builtin_amdgcn_readlane(pv, i);
            const int32_t nci = __builtin_amdgcn_readlane(nc, i);
            const int32_t nbzi = __builtin_amdgcn_readlane(nbz, i);
            const int32_t offi = __builtin_amdgcn_readlane(off, i);
            const int di = decide_seq(x, s0 + i, vi, pvi, nci, nbzi, offi);
            if (di >= 0) {
                if (lane == 0) {
                    for (int c = 0; c < ncon; ++c) {
                        const int32_t wc = c == 0 ? __builtin_amdgcn_readlane(w0, i)
                                                  : (c == 1 ? __builtin_amdgcn_readlane(w1, i) : vw_of(a, vi, c));
                        x.pw[di * ncon + c] += wc;
                        x.pw[pvi * ncon + c] -= wc;
                    }
                    a.moves[nout + moved] = make_int2(vi, di);
                    a.part[vi] = di;
                }
                ++moved;
            }
            wave_sync();
        }
        return moved;
    }
    // cut-optimisation, <= kCK candidates per lane, <= 2 constraints
    int32_t cq[kCK], b0[kCK], b1[kCK], f0[kCK], f1[kCK];
#pragma unroll
    for (int t = 0; t < kCK; ++t) {
        cq[t] = -3;
        b0[t] = b1[t] = 0;
        f0[t] = INT32_MIN;
        f1[t] = INT32_MAX;
        if (act && t < nc) {
            const int32_t q = x.w.cl[s * kCL + t];
            cq[t] = q;
            b0[t] = x.pw[q * ncon];
            f0[t] = x.fm[q * ncon];
            if (ncon > 1) {
                b1[t] = x.pw[q * ncon + 1];
                f1[t] = x.fm[q * ncon + 1];
            }
        }
    }
    auto choose = [&](const int32_t* d0, const int32_t* d1) -> int {
#pragma unroll
        for (int t = 0; t < kCK; ++t)
            if (cq[t] >= 0 && (long long)b0[t] + d0[t] + w0 <= f0[t] && (long long)b1[t] + d1[t] + w1 <= f1[t])
                return cq[t];
        return kNoDest;
    };
    const int32_t z[kCK] = {0, 0, 0, 0};
    int d = act ? choose(z, z) : kNoDest;
    // box certification against the chunk's tentative flows
    if (d >= 0) {
        atomicAdd(&x.A[d * ncon], w0);
        atomicAdd(&x.R[pv * ncon], w0);
        if (ncon > 1) {
            atomicAdd(&x.A[d * ncon + 1], w1);
            atomicAdd(&x.R[pv * ncon + 1], w1);
        }
    }
    wave_sync();
    bool ok = true;
#pragma unroll
    for (int t = 0; t < kCK; ++t) {
        if (!ok || cq[t] < 0) continue;
        const int32_t q = cq[t];
        if (q == d) {
            const long long a0 = x.A[q * ncon] - w0;
            const long long a1 = ncon > 1 ? x.A[q * ncon + 1] - w1 : 0;
            if (b0[t] + a0 + w0 > f0[t] || b1[t] + a1 + w1 > f1[t]) ok = false;
            break;
        }
        const long long r0 = x.R[q * ncon];
        const long long r1 = ncon > 1 ? x.R[q * ncon + 1] : 0;
        if (b0[t] - r0 + w0 <= f0[t] && b1[t] - r1 + w1 <= f1[t]) ok = false;
    }
    wave_sync();
    if (d >= 0) {
        x.A[d * ncon] = 0;
        x.R[pv * ncon] = 0;
        if (ncon > 1) {
            x.A[d * ncon + 1] = 0;
            x.R[pv * ncon + 1] = 0;
        }
    }
    wave_sync();
    if (__ballot(act && !ok)) {
        for (int iter = 0; iter <= 64; ++iter) {
            int32_t d0[kCK] = {0, 0, 0, 0}, d1[kCK] = {0, 0, 0, 0};
            uint64_t mm = __ballot(d >= 0);
            while (mm) {
                const int j = __ffsll((unsigned long long)mm) - 1;
                mm &= mm - 1;
                const int32_t dj = __builtin_amdgcn_readlane(d, j);
                const int32_t pj = __builtin_amdgcn_readlane(pv, j);
                const int32_t wj0 = __builtin_amdgcn_readlane(w0, j);
                const int32_t wj1 = __builtin_amdgcn_readlane(w1, j);
                if (j < lane) {
#pragma unroll
                    for (int t = 0; t < kCK; ++t) {
                        if (cq[t] == dj) {
                            d0[t] += wj0;
                            d1[t] += wj1;
                        }
                        if (cq[t] == pj) {
                            d0[t] -= wj0;
                            d1[t] -= wj1;
                        }
                    }
                }
            }
            const int nd = act ? choose(d0, d1) : kNoDest;
            const bool changed = nd != d;
            d = nd;
            if (__ballot(changed) == 0) break;
        }
    }
    if (d >= 0) {
        atomicAdd(&x.pw[d * ncon], w0);
        atomicSub(&x.pw[pv * ncon], w0);
        if (ncon > 1) {
            atomicAdd(&x.pw[d * ncon + 1], w1);
            atomicSub(&x.pw[pv * ncon + 1], w1);
        }
    }
    const uint64_t mv = __ballot(d >= 0);
    if (d >= 0) {
        const int pos = nout + __popcll(mv & ((1ull << lane) - 1ull));
        a.moves[pos] = make_int2(v, d);
        a.part[v] = d;
    }
    wave_sync();
    return __popcll(mv);
}

__global__ __launch_bounds__(kAT) void k_admit(AdmArgs a) {
    extern __shared__ int32_t s_dyn[];
    __shared__ int s_w[kAT / 64];
    __shared__ int s_nout;
    if (*a.done) return;
    const int tid = threadIdx.x;
    const int kc = a.k * a.ncon;
    const bool in_lds = kc <= kLdsKC;
    int32_t* base = in_lds ? s_dyn : a.gscratch;
    int32_t* pw = in_lds ? base : a.pw;
    int32_t* fm = base + kc;
    int32_t* om = base + 2 * kc;
    int32_t* A = base + 3 * kc;
    int32_t* R = in_lds ? base + 4 * kc : base;   // global layout: R, fm, om, A
    for (int i = tid; i < kc; i += kAT) {
        if (in_lds) pw[i] = a.pw[i];
        fm[i] = clamp32(a.fmax[i]);
        om[i] = clamp32(a.omax[i]);
        A[i] = 0;
        R[i] = 0;
    }
    // union region (16-byte aligned): tile counts, then the mover window
    int32_t* uni = in_lds ? s_dyn + ((5 * kc + 3) & ~3) : s_dyn;
    // 1. ordered gather of the per-tile records into the dense list a.mov
    int nmov = 0;
    for (int tw = 0; tw < a.ntiles; tw += kCntWin) {
        const int nt = min(kCntWin, a.ntiles - tw);
        for (int i = tid; i < nt; i += kAT) uni[i] = a.tile_cnt[tw + i];
        __syncthreads();
        const int per = (nt + kAT - 1) / kAT;
        const int t0 = min(tid * per, nt), t1 = min(t0 + per, nt);
        int cnt = 0;
        for (int t = t0; t < t1; ++t) cnt += uni[t];
        int wtot;
        int off = nmov + block_excl_scan_int(cnt, s_w, &wtot);
        for (int t = t0; t < t1; ++t) {
            const int c = uni[t];
            const int4* src = a.tile_rec + (int64_t)(tw + t) * kTile;
            for (int i = 0; i < c; ++i) a.mov[off + i] = src[i];
            off += c;
        }
        nmov += wtot;
        __syncthreads();
    }
    __threadfence();
    __syncthreads();
    // 2. windows of kWin movers: stage in LDS, resolve with wave 0
    Win w;
    w.rec = (int4*)uni;
    w.off = uni + 4 * kWin;
    w.w0 = w.off + kWin;
    w.w1 = w.w0 + kWin;
    w.cl = w.w1 + kWin;
    Ctx x{&a, pw, fm, om, A, R, w};
    if (tid == 0) s_nout = 0;
    for (int W0 = 0; W0 < nmov; W0 += kWin) {
        const int i = W0 + tid;
        if (i < nmov) {
            const int4 r = a.mov[i];
            const int32_t off = a.xadj[r.x];
            w.rec[tid] = r;
            w.off[tid] = off;
            w.w0[tid] = vw_of(a, r.x, 0);
            w.w1[tid] = a.ncon > 1 ? vw_of(a, r.x, 1) : 0;
            const int lim = r.z < kCL ? r.z : kCL;
            for (int t = 0; t < lim; ++t) w.cl[tid * kCL + t] = a.pool[off + t];
        }
        __syncthreads();
        if ((tid >> 6) == 0) {
            const int wn = min(kWin, nmov - W0);
            int nout = s_nout;
            for (int s0 = 0; s0 < wn; s0 += 64) nout += resolve_chunk(x, s0, min(64, wn - s0), nout);
            if (tid == 0) s_nout = nout;
        }
        __syncthreads();
    }
    __threadfence();
    __syncthreads();
    if (in_lds)
        for (int i = tid; i < kc; i += kAT) a.pw[i] = pw[i];
    if (tid == 0) {
        *a.nmoves = s_nout;
        *a.passes += 1;
        if (a.mode == 1 && s_nout == 0) *a.done = 1;
    }
}

}  // namespace

void admit_pass(const DevGraph& g, int32_t k, int mode, int ntiles, RefineBufs& b, int32_t* part, hipStream_t s) {
    AdmArgs aa{mode, k, g.ncon, ntiles, b.tile_cnt, b.tile_rec, b.mov, b.pool, g.xadj, g.vwgt, b.fmax, b.omax,
               b.pw, part, b.moves, b.nmoves, b.done, b.passes, b.adm_scratch};
    const int kc = k * g.ncon;
    const size_t win = (size_t)kWin * (4 + 1 + 1 + 1 + kCL) * sizeof(int32_t);
    const size_t cntb = (size_t)kCntWin * sizeof(int32_t);
    const size_t uni = win > cntb ? win : cntb;
    const size_t lds = (kc <= kLdsKC ? (size_t)((5 * kc + 3) & ~3) * sizeof(int32_t) : 0) + uni;
    hipLaunchKernelGGL(k_admit, dim3(1), dim3(kAT), lds, s, aa);
    RK_HIP(hipGetLastError());
}

}  // namespace rk
