// refine.hip -- one synchronous greedy round on the GPU (refine/greedy.rs:36-193).
//
// A round of the reference visits the boundary vertices B in ascending order.
// Gains use the round-start partition, but feasibility (and the repair pass's
// "source overweight" test) use a running part-weight vector pw that is
// updated after every recorded move; the moves are applied after the scan.
// The GPU reproduces that bit-exactly in three steps:
//
//  1. k_gain (G lanes per vertex) / k_gain_heavy (a block per vertex of
//     degree > 4G): boundary test and the neighbour-part gain table
//     ed[p] = sum of edge weights to part p (compute_move_gain,
//     greedy.rs:197-218, gives gain(v,pv,p) = ed[p] - ed[pv]).  Every vertex
//     that could move gets a record and its candidate parts in admission
//     order -- descending gain, ties to the lowest part -- which is the order
//     in which the reference's strict `gain > best_gain` scan over p = 0..k-1
//     picks the first feasible part.  Parts with ed == 0 (all non-adjacent
//     parts) share one gain value and are kept implicit ("zero group").
//  2. k_compact: records of all tiles, in vertex order, into one dense list.
//  3. k_admit (one workgroup): walks the list in order.  A chunk of movers
//     decides against the chunk-start pw; the chunk's tentative in/out flows
//     A, R give a box [pw - R, pw + A] that contains every exact prefix
//     weight; a mover whose decision is the same at both corners (feasibility
//     and the overweight test are monotone) is certified.  The certified
//     prefix is committed; the first uncertified mover is then decided
//     against the exact running pw, so every step makes progress and the
//     result equals the sequential scan.
//
// Thresholds: `w as f64 <= tpw*ub` (greedy.rs:87,144) and `> tpw*ub`
// (greedy.rs:70) are evaluated as integer compares against floor(tpw*ub),
// computed on the host per level (fmax for feasibility, omax for overweight).
#include <climits>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "profiler.hpp"

namespace rk {

namespace {

constexpr int kGateOff = -2;   // repair pass: source part not overweight
constexpr int kNoDest = -1;    // no feasible destination
constexpr int kAdmLds = 4096;  // k*ncon kept in LDS by the admission kernel
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
    const int32_t* done;
    int4* vrec;                 // optional per-vertex record (gain-table export)
    int32_t* g0_out;            // optional zero-group gain per vertex
    int32_t* gains;             // optional candidate gains beside pool
};

__device__ inline bool gain_qualifies(int mode, int32_t g) {
    // cut-optimisation: best_gain starts at 0 with strict > (greedy.rs:135,150);
    // repair: best_gain starts at i32::MIN with strict > (greedy.rs:78,93).
    return mode == 1 ? g > 0 : g != INT32_MIN;
}

__device__ inline bool part_overweight(const GainArgs& a, int32_t p) {
    for (int c = 0; c < a.ncon; ++c)
        if ((long long)a.pw[(int64_t)p * a.ncon + c] > a.omax[(int64_t)p * a.ncon + c]) return true;
    return false;
}

__device__ inline void lds_hash_add(int* hk, int* hv, int S, int32_t key, int32_t w) {
    int s = key & (S - 1);
    for (int probe = 0; probe < S; ++probe) {
        int prev = atomicCAS(&hk[s], -1, key);
        if (prev == -1 || prev == key) {
            atomicAdd(&hv[s], w);
            return;
        }
        s = (s + 1) & (S - 1);
    }
}

template <int G>
__global__ __launch_bounds__(kBlock) void k_gain(GainArgs a) {
    constexpr int NG = kBlock / G;
    constexpr int S = 4 * G;
    __shared__ int s_key[NG * S];
    __shared__ int s_val[NG * S];
    __shared__ int4 s_rec[kTile];
    __shared__ int s_tmp[kBlock / 64];
    if (*a.done) return;
    const int grp = threadIdx.x / G;
    const int gl = group_lane<G>();
    int* hk = s_key + grp * S;
    int* hv = s_val + grp * S;
    for (int s = gl; s < S; s += G) {
        hk[s] = -1;
        hv[s] = 0;
    }
    wave_sync();
    const int64_t tile0 = (int64_t)blockIdx.x * kTile;
    for (int it = 0; it < kTile / NG; ++it) {
        const int vi = it * NG + grp;
        const int64_t v = tile0 + vi;
        int4 rec = make_int4(-1, 0, 0, 0);
        if (v < a.n) {
            const int32_t beg = a.xadj[v], end = a.xadj[v + 1];
            const int32_t pv = a.part[v];
            if (end - beg > S) {
                rec = a.heavy_rec[v];
            } else {
                uint32_t internal = 0;
                bool ext = false;
                for (int32_t j = beg + gl; j < end; j += G) {
                    const int32_t u = a.adjncy[j];
                    const int32_t w = a.adjwgt ? a.adjwgt[j] : 1;
                    const int32_t pu = a.part[u];
                    if (pu == pv) {
                        internal += (uint32_t)w;
                    } else {
                        ext = true;
                        lds_hash_add(hk, hv, S, pu, w);
                    }
                }
                wave_sync();
                const int32_t ed_own = (int32_t)group_sum<G>(internal);
                const bool boundary = group_ballot<G>(ext) != 0;
                if (boundary) {
                    const bool gate = a.mode == 1 || part_overweight(a, pv);
                    if (gate) {
                        const int32_t g0 = wsub(0, ed_own);
                        int ncand = 0, nbefore = 0;
                        for (int s = gl; s < S; s += G) {
                            const int32_t key = hk[s];
                            const int32_t val = hv[s];
                            if (key >= 0 && val != 0) {
                                const int32_t g = wsub(val, ed_own);
                                if (gain_qualifies(a.mode, g)) {
                                    ++ncand;
                                    if (g > g0) ++nbefore;
                                }
                            }
                        }
                        ncand = group_sum<G>(ncand);
                        nbefore = group_sum<G>(nbefore);
                        const bool zflag = gain_qualifies(a.mode, g0);
                        if (ncand > 0 || zflag) {
                            for (int s = gl; s < S; s += G) {
                                const int32_t key = hk[s];
                                const int32_t val = hv[s];
                                if (key < 0 || val == 0) continue;
                                const int32_t g = wsub(val, ed_own);
                                if (!gain_qualifies(a.mode, g)) continue;
                                int rank = 0;
                                for (int s2 = 0; s2 < S; ++s2) {
                                    const int32_t k2 = hk[s2];
                                    const int32_t v2 = hv[s2];
                                    if (k2 < 0 || v2 == 0) continue;
                                    const int32_t g2 = wsub(v2, ed_own);
                                    if (!gain_qualifies(a.mode, g2)) continue;
                                    if (g2 > g || (g2 == g && k2 < key)) ++rank;
                                }
                                a.pool[beg + rank] = key;
                                if (a.gains) a.gains[beg + rank] = g;
                            }
                            if (a.g0_out && gl == 0) a.g0_out[v] = g0;
                            rec = make_int4((int)v, pv, ncand,
                                            (int)((uint32_t)nbefore | (zflag ? 0x80000000u : 0u)));
                        }
                    }
                    wave_sync();
                    for (int s = gl; s < S; s += G) {
                        hk[s] = -1;
                        hv[s] = 0;
                    }
                    wave_sync();
                }
            }
        }
        if (gl == 0) {
            s_rec[vi] = rec;
            if (a.vrec && v < a.n) a.vrec[v] = rec;
        }
    }
    __syncthreads();
    const int4 r = s_rec[threadIdx.x];
    const bool f = r.x >= 0;
    int tot;
    const int pos = block_flag_scan(f, s_tmp, &tot);
    if (f) a.tile_rec[(int64_t)blockIdx.x * kTile + pos] = r;
    if (threadIdx.x == 0) a.tile_cnt[blockIdx.x] = tot;
}

// One block per heavy vertex; dense neighbour-part histogram of k entries in
// LDS (or in a per-block global scratch row when k > kHeavyLds).
__global__ __launch_bounds__(kBlock) void k_gain_heavy(GainArgs a, const int32_t* __restrict__ heavy_list,
                                                       int32_t nheavy, int4* __restrict__ heavy_rec,
                                                       int32_t* __restrict__ gscratch) {
    extern __shared__ int32_t s_dyn[];
    __shared__ uint32_t s_int;
    __shared__ int s_ext, s_nc, s_nb;
    if (*a.done) return;
    const int k = a.k;
    int32_t* hist;
    int2* cands;
    if (k <= kHeavyLds) {
        hist = s_dyn;
        cands = (int2*)(s_dyn + k + (k & 1));
    } else {
        hist = gscratch + (int64_t)blockIdx.x * 3 * k;
        cands = (int2*)(hist + k + (k & 1));
    }
    for (int hi = blockIdx.x; hi < nheavy; hi += gridDim.x) {
        const int32_t v = heavy_list[hi];
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
        const bool gate = a.mode == 1 || part_overweight(a, pv);
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

__global__ __launch_bounds__(kBlock) void k_heavy_list(int32_t n, const int32_t* __restrict__ xadj,
                                                       int32_t thresh, int32_t* __restrict__ list,
                                                       int32_t* __restrict__ count) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool h = v < n && (xadj[v + 1] - xadj[v]) > thresh;
    uint64_t b = __ballot(h);
    int base = 0;
    if (__lane_id() == 0 && b) base = atomicAdd(count, __popcll(b));
    base = __shfl(base, 0, 64);
    if (h) list[base + __popcll(b & ((1ull << __lane_id()) - 1ull))] = (int32_t)v;
}

__global__ __launch_bounds__(kBlock) void k_compact(const int4* __restrict__ tile_rec,
                                                    const int32_t* __restrict__ tile_cnt,
                                                    const int32_t* __restrict__ tile_off,
                                                    int4* __restrict__ mov, const int32_t* __restrict__ done) {
    if (*done) return;
    const int cnt = tile_cnt[blockIdx.x];
    const int off = tile_off[blockIdx.x];
    for (int i = threadIdx.x; i < cnt; i += kBlock) mov[off + i] = tile_rec[(int64_t)blockIdx.x * kTile + i];
}

struct AdmitArgs {
    int32_t mode, k, ncon;
    const int32_t* nmov;
    const int4* mov;
    const int32_t* pool;
    const int32_t* xadj;
    const int32_t* vwgt;
    const long long* fmax;
    const long long* omax;
    int32_t* pw;
    int32_t* gA;
    int32_t* gR;
    int32_t* part;
    int2* moves;
    int32_t* nmoves;
    int32_t* done;
    int32_t* passes;
};

// pw at one corner of the certification box.  dir 0: exact; +1: pw + A minus
// the mover's own inflow; -1: pw - R minus the mover's own outflow.
struct View {
    const int32_t* pw;
    const int32_t* X;
    int dir, d, pv;
    __device__ long long get(int q, int c, int ncon, int32_t vwc) const {
        const int idx = q * ncon + c;
        long long x = pw[idx];
        if (dir > 0) {
            long long t = X[idx];
            if (q == d) t -= vwc;
            x += t;
        } else if (dir < 0) {
            long long t = X[idx];
            if (q == pv && d >= 0) t -= vwc;
            x -= t;
        }
        return x;
    }
};

struct Mover {
    int32_t v, pv, nc, nbz;
    const int32_t* cl;   // candidate parts in admission order
    const int32_t* vw;   // vwgt row or nullptr (unit)
    __device__ int32_t w(int c) const { return vw ? vw[c] : 1; }
};

__device__ inline bool feasible(const AdmitArgs& a, const View& vw, const Mover& m, int q) {
    // greedy.rs:85-88 / 142-145: all c: pw[q][c] + vw[c] <= tpw*ub
    for (int c = 0; c < a.ncon; ++c) {
        const int32_t wc = m.w(c);
        if (vw.get(q, c, a.ncon, wc) + wc > a.fmax[(int64_t)q * a.ncon + c]) return false;
    }
    return true;
}

__device__ inline bool overweight(const AdmitArgs& a, const View& vw, const Mover& m) {
    // greedy.rs:69-71: any c: pw[pv][c] > tpw*ub
    for (int c = 0; c < a.ncon; ++c)
        if (vw.get(m.pv, c, a.ncon, m.w(c)) > a.omax[(int64_t)m.pv * a.ncon + c]) return true;
    return false;
}

// Visit the mover's candidates in admission order until f returns true.
template <typename F>
__device__ inline void walk(const AdmitArgs& a, const Mover& m, F&& f) {
    const int nbefore = m.nbz & 0x7fffffff;
    const bool zflag = (uint32_t)m.nbz >> 31;
    for (int t = 0; t < nbefore; ++t)
        if (f(m.cl[t])) return;
    if (zflag) {
        for (int q = 0; q < a.k; ++q) {
            if (q == m.pv) continue;
            bool expl = false;
            for (int t = 0; t < m.nc; ++t)
                if (m.cl[t] == q) {
                    expl = true;
                    break;
                }
            if (expl) continue;
            if (f(q)) return;
        }
    }
    for (int t = nbefore; t < m.nc; ++t)
        if (f(m.cl[t])) return;
}

__device__ inline int decide(const AdmitArgs& a, const View& v0, const Mover& m) {
    if (a.mode == 0 && !overweight(a, v0, m)) return kGateOff;
    int d = kNoDest;
    walk(a, m, [&](int q) {
        if (feasible(a, v0, m, q)) {
            d = q;
            return true;
        }
        return false;
    });
    return d;
}

__device__ inline bool certify(const AdmitArgs& a, const View& lo, const View& hi, const Mover& m, int d) {
    if (a.mode == 0) {
        if (d == kGateOff) return !overweight(a, hi, m);
        if (!overweight(a, lo, m)) return false;
    }
    bool ok = true;
    walk(a, m, [&](int q) {
        if (q == d) {
            ok = feasible(a, hi, m, q);
            return true;
        }
        if (feasible(a, lo, m, q)) {
            ok = false;
            return true;
        }
        return false;
    });
    return ok;
}

__global__ __launch_bounds__(kBlock) void k_admit(AdmitArgs a) {
    extern __shared__ int32_t s_dyn[];
    __shared__ int s_fail, s_nout;
    __shared__ int s_tmp[kBlock / 64];
    if (*a.done) return;
    const int kc = a.k * a.ncon;
    int32_t *pw, *A, *R;
    const bool in_lds = kc <= kAdmLds;
    if (in_lds) {
        pw = s_dyn;
        A = s_dyn + kc;
        R = s_dyn + 2 * kc;
        for (int i = threadIdx.x; i < kc; i += kBlock) {
            pw[i] = a.pw[i];
            A[i] = 0;
            R[i] = 0;
        }
    } else {
        pw = a.pw;
        A = a.gA;
        R = a.gR;
    }
    if (threadIdx.x == 0) s_nout = 0;
    __syncthreads();
    const int nmov = *a.nmov;
    for (int cb = 0; cb < nmov; cb += kBlock) {
        const int i = cb + threadIdx.x;
        const bool active = i < nmov;
        Mover m{};
        if (active) {
            const int4 r = a.mov[i];
            m.v = r.x;
            m.pv = r.y;
            m.nc = r.z;
            m.nbz = r.w;
            m.cl = a.pool + a.xadj[m.v];
            m.vw = a.vwgt ? a.vwgt + (int64_t)m.v * a.ncon : nullptr;
        }
        int start = cb;
        for (;;) {
            const bool live = active && i >= start;
            int d = kNoDest;
            if (live) {
                View v0{pw, nullptr, 0, -1, m.pv};
                d = decide(a, v0, m);
                if (d >= 0)
                    for (int c = 0; c < a.ncon; ++c) {
                        atomicAdd(&A[d * a.ncon + c], m.w(c));
                        atomicAdd(&R[m.pv * a.ncon + c], m.w(c));
                    }
            }
            if (threadIdx.x == 0) s_fail = INT32_MAX;
            __syncthreads();
            if (live && i > start) {
                View lo{pw, R, -1, d, m.pv};
                View hi{pw, A, +1, d, m.pv};
                if (!certify(a, lo, hi, m, d)) atomicMin(&s_fail, i);
            }
            __syncthreads();
            const int f = s_fail;
            const bool mv = live && i < f && d >= 0;
            int tot;
            const int pos = block_flag_scan(mv, s_tmp, &tot);
            const int base = s_nout;
            if (live && d >= 0)
                for (int c = 0; c < a.ncon; ++c) {
                    A[d * a.ncon + c] = 0;
                    R[m.pv * a.ncon + c] = 0;
                }
            __syncthreads();
            if (mv) {
                a.moves[base + pos] = make_int2(m.v, d);
                a.part[m.v] = d;
                for (int c = 0; c < a.ncon; ++c) {
                    atomicAdd(&pw[d * a.ncon + c], m.w(c));
                    atomicSub(&pw[m.pv * a.ncon + c], m.w(c));
                }
            }
            if (threadIdx.x == 0) s_nout = base + tot;
            __syncthreads();
            if (f == INT32_MAX) break;
            start = f;
        }
    }
    if (in_lds)
        for (int i = threadIdx.x; i < kc; i += kBlock) a.pw[i] = pw[i];
    if (threadIdx.x == 0) {
        *a.nmoves = s_nout;
        *a.passes += 1;
        if (a.mode == 1 && s_nout == 0) *a.done = 1;
    }
}

void launch_gain(const GainArgs& ga, int G, int nheavy, RefineBufs& b, hipStream_t s) {
    const int ntiles = (ga.n + kTile - 1) / kTile;
    const int k = ga.k;
    if (nheavy > 0) {
        const size_t lds = k <= kHeavyLds ? (size_t)(k + (k & 1) + 2 * k) * sizeof(int32_t) : 0;
        const int cap = k <= kHeavyLds ? 2048 : kHeavyGlobalBlocks;
        const int grid = nheavy < cap ? nheavy : cap;
        hipLaunchKernelGGL(k_gain_heavy, dim3(grid), dim3(kBlock), lds, s, ga, b.heavy_list, nheavy,
                           b.heavy_rec, b.heavy_scratch);
    }
    switch (G) {
        case 4: hipLaunchKernelGGL(k_gain<4>, dim3(ntiles), dim3(kBlock), 0, s, ga); break;
        case 8: hipLaunchKernelGGL(k_gain<8>, dim3(ntiles), dim3(kBlock), 0, s, ga); break;
        case 16: hipLaunchKernelGGL(k_gain<16>, dim3(ntiles), dim3(kBlock), 0, s, ga); break;
        case 32: hipLaunchKernelGGL(k_gain<32>, dim3(ntiles), dim3(kBlock), 0, s, ga); break;
        default: hipLaunchKernelGGL(k_gain<64>, dim3(ntiles), dim3(kBlock), 0, s, ga); break;
    }
}

}  // namespace

void gain_table(const DevGraph& g, int32_t k, int G, int nheavy, RefineBufs& b, const int32_t* part, int4* vrec,
                int32_t* g0, int32_t* gains, hipStream_t s) {
    if (g.n <= 0) return;
    // Every part overweight (pw 0 > omax = INT64_MIN) so that every boundary
    // vertex lists all its adjacent parts with a nonzero edge-weight sum.
    const size_t kc = (size_t)k * g.ncon;
    RK_HIP(hipMemsetAsync(b.pw, 0, kc * sizeof(int32_t), s));
    std::vector<long long> om(kc, LLONG_MIN);
    RK_HIP(hipMemcpyAsync(b.omax, om.data(), kc * sizeof(long long), hipMemcpyHostToDevice, s));
    RK_HIP(hipMemsetAsync(b.done, 0, sizeof(int32_t), s));
    GainArgs ga{g.n, g.ncon, k, 0, g.xadj, g.adjncy, g.adjwgt, g.vwgt, part, b.pw, b.omax,
                b.heavy_rec, b.pool, b.tile_rec, b.tile_cnt, b.done, vrec, g0, gains};
    launch_gain(ga, G, nheavy, b, s);
    RK_HIP(hipStreamSynchronize(s));
    RK_HIP(hipGetLastError());
}

int refine_setup_level(const DevGraph& g, int G, RefineBufs& b, hipStream_t s) {
    RK_HIP(hipMemsetAsync(b.heavy_cnt, 0, sizeof(int32_t), s));
    if (g.n > 0)
        hipLaunchKernelGGL(k_heavy_list, dim3((g.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, g.n, g.xadj,
                           4 * G, b.heavy_list, b.heavy_cnt);
    RK_HIP(hipGetLastError());
    int32_t nh = 0;
    RK_HIP(hipMemcpyAsync(&nh, b.heavy_cnt, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RK_HIP(hipStreamSynchronize(s));
    return nh;
}

void refine_pass(const DevGraph& g, int32_t k, int mode, int G, int nheavy, RefineBufs& b, int32_t* part,
                 hipStream_t s) {
    if (g.n <= 0) return;
    const int ntiles = (g.n + kTile - 1) / kTile;
    GainArgs ga{g.n, g.ncon, k, mode, g.xadj, g.adjncy, g.adjwgt, g.vwgt, part, b.pw, b.omax,
                b.heavy_rec, b.pool, b.tile_rec, b.tile_cnt, b.done, nullptr, nullptr, nullptr};
    {
        // algorithmic bytes: xadj, adjncy, adjwgt, part[adjncy] gather, part[v]
        const double bytes = 4.0 * (g.n + 1) + 4.0 * g.n + (g.adjwgt ? 12.0 : 8.0) * g.m;
        ProfScope ps(kProfGain, bytes, g.n, s);
        launch_gain(ga, G, nheavy, b, s);
    }
    {
        ProfScope ps(kProfCompact, 12.0 * ntiles, g.n, s);
        exclusive_scan_i32(b.tile_cnt, b.tile_off, ntiles, b.nmov, b.scan_tmp, s);
        hipLaunchKernelGGL(k_compact, dim3(ntiles), dim3(kBlock), 0, s, b.tile_rec, b.tile_cnt, b.tile_off,
                           b.mov, b.done);
    }
    AdmitArgs aa{mode, k, g.ncon, b.nmov, b.mov, b.pool, g.xadj, g.vwgt, b.fmax, b.omax, b.pw, b.gA, b.gR,
                 part, b.moves, b.nmoves, b.done, b.passes};
    const int kc = k * g.ncon;
    if (kc > kAdmLds) {
        RK_HIP(hipMemsetAsync(b.gA, 0, (size_t)kc * sizeof(int32_t), s));
        RK_HIP(hipMemsetAsync(b.gR, 0, (size_t)kc * sizeof(int32_t), s));
    }
    const size_t lds = kc <= kAdmLds ? (size_t)3 * kc * sizeof(int32_t) : 0;
    {
        ProfScope ps(kProfAdmit, 0.0, g.n, s);
        hipLaunchKernelGGL(k_admit, dim3(1), dim3(kBlock), lds, s, aa);
    }
    RK_HIP(hipGetLastError());
}

}  // namespace rk
