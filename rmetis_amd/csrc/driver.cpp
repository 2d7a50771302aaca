// driver.cpp -- multilevel k-way scheduler on the host (partition/kway.rs:20-162).
//
// Per attempt: coarsen on the device until the level is small or stalls
// (coarsen/mod.rs:36-71), grow an initial partition of the coarsest level on
// the host (initial/grow.rs:12-103), then refine the coarsest level and every
// finer level after projecting (kway.rs:73-92).  Keeps the strictly best cut
// over the ncuts attempts (kway.rs:96-98).
#include "driver.hpp"
#include "tuning.hpp"

#include "profiler.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <deque>
#include <random>

namespace rk {

Options parse_options(const int32_t* a) {
    Options o;
    if (!a || a[0] == -1) return o;                 // c_api.rs:211-222
    o.ctype_shem = a[2] == 0 ? 0 : 1;               // types.rs:192-194
    o.iptype = a[3] == 1 ? 1 : (a[3] == 4 ? 4 : 0); // types.rs:195-199
    o.rtype_greedy = a[4] == 1 ? 1 : 0;             // types.rs:200-202
    o.objtype = a[1] == 1 ? 1 : 0;
    o.no2hop = a[9] != 0 ? 1 : 0;
    if (a[7] > 0) o.ncuts = a[7];
    if (a[6] > 0) o.niter = a[6];
    o.seed = a[8] > 0 ? (uint64_t)a[8] : 0;         // types.rs:215: max(0)
    o.minconn = a[10] != 0;
    o.contig = a[11] != 0;
    if (a[16] > 0) o.ufactor = a[16];
    o.numbering = (uint8_t)a[17];
    o.dbglvl = a[5];
    return o;
}

void default_options_raw(int32_t* arr) {
    // types.rs:227-245 with Options::default() (types.rs:77-95)
    for (int i = 0; i < kNOptions; ++i) arr[i] = -1;
    arr[2] = 1;    // CTYPE  Shem
    arr[3] = 0;    // IPTYPE Grow
    arr[4] = 0;    // RTYPE  Fm
    arr[1] = 0;    // OBJTYPE Cut
    arr[9] = 0;    // NO2HOP
    arr[7] = 1;    // NCUTS
    arr[6] = 10;   // NITER
    arr[8] = 0;    // SEED
    arr[10] = 0;   // MINCONN
    arr[11] = 0;   // CONTIG
    arr[16] = 30;  // UFACTOR
    arr[17] = 0;   // NUMBERING
    arr[5] = 0;    // DBGLVL
}

namespace {
std::atomic<int32_t> g_debug_flags{0};
}  // namespace

void set_debug_flags(int32_t flags) { g_debug_flags.store(flags, std::memory_order_relaxed); }
int32_t debug_flags() { return g_debug_flags.load(std::memory_order_relaxed); }

namespace {

// For integer w and real t: (w as f64 <= t) <=> w <= floor(t), and
// (w as f64 > t) <=> w > floor(t).  NaN compares false both ways.
long long floor_threshold(double t, bool feasibility) {
    if (std::isnan(t)) return feasibility ? LLONG_MIN : LLONG_MAX;
    if (t >= 9.0e18) return LLONG_MAX;
    if (t <= -9.0e18) return LLONG_MIN;
    return (long long)std::floor(t);
}

}  // namespace

void level_thresholds(int32_t k, int32_t ncon, const float* tpwgts, const int32_t* total, const double* ub,
                      long long* fmax, long long* omax) {
    for (int32_t p = 0; p < k; ++p)
        for (int32_t c = 0; c < ncon; ++c) {
            const size_t i = (size_t)p * ncon + c;
            const double f = tpwgts ? (double)tpwgts[i] : 1.0 / (double)k;   // graph/mod.rs:401-405
            const double target = f * (double)total[c];                        // graph/mod.rs:406
            const double t = target * ub[c];                                   // greedy.rs:70,87,144
            fmax[i] = floor_threshold(t, true);
            omax[i] = floor_threshold(t, false);
        }
}

void thresholds_from_targets(int32_t k, int32_t ncon, const double* tpw, const double* ub, long long* fmax,
                             long long* omax) {
    for (int32_t p = 0; p < k; ++p)
        for (int32_t c = 0; c < ncon; ++c) {
            const size_t i = (size_t)p * ncon + c;
            const double t = tpw[i] * ub[c];
            fmax[i] = floor_threshold(t, true);
            omax[i] = floor_threshold(t, false);
        }
}

namespace {

// The per-call work buffers, laid out through A: the arena when allocating,
// a byte counter when sizing (buffers_bytes), so that both see one list.
template <class A>
Buffers layout_buffers(A& a, int32_t n, int32_t m, int32_t k, int32_t ncon, int64_t ntiles, int64_t kc) {
    Buffers B;
    RefineBufs& r = B.rb;
    r.tile_rec = a.template get<int4>(ntiles * kTile);
    r.tile_cnt = a.template get<int32_t>(ntiles);
    r.tile_off = a.template get<int32_t>(ntiles);
    r.tile_bits = (uint32_t*)a.template get<int32_t>(2 * ((ntiles + 31) / 32));
    r.tile_dirty = r.tile_bits + (ntiles + 31) / 32;
    r.mov = a.template get<int4>(n);
    r.pool = a.template get<int32_t>(m);
    r.heavy_rec = a.template get<int4>(n);
    r.heavy_scratch = k > 4096 ? a.template get<int32_t>((int64_t)kHeavyGlobalBlocks * (3 * (int64_t)k + 2)) : nullptr;
    r.moves = a.template get<int2>(n);
    r.pw = a.template get<int32_t>(kc);
    r.fmax = a.template get<long long>(kc);
    r.omax = a.template get<long long>(kc);
    r.adm_scratch = a.template get<int32_t>(4 * kc);
    int32_t* cnt = a.template get<int32_t>(64);
    r.heavy_cnt = cnt + 0;
    r.nmov = cnt + 1;
    r.nmoves = cnt + 2;
    r.done = cnt + 3;
    r.passes = cnt + 4;
    r.full = cnt + 5;
    r.scan_tmp = a.alloc(scan_temp_bytes(std::max<int64_t>(ntiles, n)));
    CoarsenBufs& c = B.cb;
    c.match = a.template get<int32_t>(n);
    c.cand = a.template get<int32_t>(n);
    c.leader = a.template get<int32_t>(n);
    c.cid = a.template get<int32_t>(n);
    c.members = a.template get<int32_t>(2 * (int64_t)n);
    c.degsum = a.template get<int32_t>(n);
    c.tmp_off = a.template get<int32_t>(n);
    c.tmp_adj = a.template get<int32_t>(m);
    c.tmp_w = a.template get<int32_t>(m);
    c.cdeg = a.template get<int32_t>(n);
    c.big_list = a.template get<int32_t>(n);
    c.hash_keys = a.template get<int32_t>(2 * (int64_t)m + n + 1);
    c.hash_first = a.template get<int32_t>(2 * (int64_t)m + n + 1);
    c.hash_sum = a.template get<int32_t>(2 * (int64_t)m + n + 1);
    c.counters = a.template get<int32_t>(16);
    c.suitor = a.template get<uint64_t>(n);
    c.scan_tmp = r.scan_tmp;
    B.cut_acc = a.template get<unsigned long long>(1);
    B.cut_out = a.template get<int32_t>(1);
    B.wsum = a.template get<long long>(2 * std::max(ncon, 1));
    return B;
}

// Sizes allocations as Arena does (256-byte granules) without memory.
struct ByteCounter {
    size_t total = 0;
    void* alloc(size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        total += bytes == 0 ? 256 : bytes;
        return nullptr;
    }
    template <typename T>
    T* get(int64_t count) {
        return (T*)alloc((size_t)(count > 0 ? count : 1) * sizeof(T));
    }
};

}  // namespace

Buffers alloc_buffers(Context& cx, int32_t n, int32_t m, int32_t k, int32_t ncon, int64_t tiles, int64_t kc_total) {
    const int64_t ntiles = tiles >= 0 ? tiles : ((int64_t)n + kTile - 1) / kTile;
    const int64_t kc = kc_total >= 0 ? kc_total : (int64_t)k * ncon;
    Buffers B = layout_buffers(cx.arena, n, m, k, ncon, ntiles, kc);
    RK_HIP(hipMemsetAsync(B.rb.tile_bits, 0, (size_t)(2 * ((ntiles + 31) / 32)) * sizeof(uint32_t), cx.stream));
    B.rb.host_small = cx.host_small();
    B.cb.host_small = B.rb.host_small;
    return B;
}

FmBufs alloc_fm_buffers(Context& cx, int32_t n, int32_t m, int32_t k, int32_t ncon) {
    Arena& a = cx.arena;
    FmBufs f;
    int64_t tree = 0, len = n;
    for (;;) {
        tree += len;
        if (len <= 1) break;
        len = (len + 63) / 64;
    }
    const int64_t kc = (int64_t)k * ncon;
    f.tree = a.get<uint64_t>(tree);
    f.tp = a.get<int32_t>(n);
    f.locked = (uint8_t*)a.get<int32_t>(((int64_t)n + 3) / 4);
    f.vrec = a.get<int4>(n);
    f.g0 = a.get<int32_t>(n);
    f.gains = a.get<int32_t>(m);
    f.pw0 = a.get<int32_t>(kc);
    f.omin = a.get<long long>(kc);
    f.ovf = a.get<int32_t>(k);
    f.err = a.get<int32_t>(1);
    fm_fill_tables(f.pw0, f.omin, kc, cx.stream);
    RK_HIP(hipMemsetAsync(f.err, 0, sizeof(int32_t), cx.stream));
    return f;
}

size_t kway_first_level_bytes(int32_t n, int32_t m, int32_t k, int32_t ncon) {
    ByteCounter c;
    const int64_t ntiles = ((int64_t)n + kTile - 1) / kTile;
    const int64_t kc = (int64_t)k * ncon;
    layout_buffers(c, n, m, k, ncon, ntiles, kc);
    c.get<int32_t>(n);                          // finest heavy-vertex list
    c.get<int32_t>(n);                          // first level: cmap
    c.get<int32_t>(n);                          //              heavy list
    c.get<int32_t>((int64_t)n * ncon);          //              coarse vwgt
    c.get<int32_t>((int64_t)n + 1);             //              coarse xadj
    c.get<int32_t>(m > 0 ? m : 1);              //              coarse adjncy
    c.get<int32_t>(m > 0 ? m : 1);              //              coarse adjwgt
    c.get<long long>(4 * kc);                   //              thresholds of two levels
    c.get<int32_t>(n);                          // finest part
    return c.total;
}

std::vector<int32_t> total_weight_checked(Context& cx, const DevGraph& g, Buffers& B, std::vector<int32_t>* maxw) {
    std::vector<int32_t> tot(g.ncon, g.n);
    if (maxw) maxw->assign(g.ncon, 1);
    if (!g.vwgt) return tot;
    std::vector<long long> s(2 * (size_t)g.ncon);
    column_sums(g, B.wsum, cx.stream, B.wsum + g.ncon);
    RK_HIP(hipMemcpyAsync(s.data(), B.wsum, sizeof(long long) * 2 * g.ncon, hipMemcpyDeviceToHost, cx.stream));
    RK_HIP(hipStreamSynchronize(cx.stream));
    for (int c = 0; c < g.ncon; ++c) {
        if (s[c] > INT32_MAX)
            throw Error(-4, "total vertex weight of a constraint exceeds the int32 range");
        tot[c] = (int32_t)s[c];
        if (maxw) (*maxw)[c] = (int32_t)s[g.ncon + c];
    }
    return tot;
}

void refine_level(Context& cx, const DevGraph& g, int32_t k, const long long* d_fmax, const long long* d_omax,
                  int niter, int32_t* d_part, int nheavy, const int32_t* heavy_list, Buffers& B, LevelStat* st,
                  bool cleared, bool pw_valid) {
    // No host synchronisation unless statistics are requested: the
    // thresholds already live on the device and the pass count is fixed
    // (a cut pass that moves nothing turns the remaining passes into no-ops).
    hipStream_t s = cx.stream;
    B.rb.fmax = const_cast<long long*>(d_fmax);
    B.rb.omax = const_cast<long long*>(d_omax);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (st) {
        RK_HIP(hipEventCreate(&e0));
        RK_HIP(hipEventCreate(&e1));
        RK_HIP(hipEventRecord(e0, s));
        RK_HIP(hipMemsetAsync(B.rb.passes, 0, sizeof(int32_t), s));
    }
    if (!pw_valid) part_weights(g, d_part, k, B.rb.pw, s, cleared);                  // greedy.rs:50
    if (!cleared) RK_HIP(hipMemsetAsync(B.rb.done, 0, sizeof(int32_t), s));
    refine_pass(g, k, 0, nheavy, heavy_list, B.rb, d_part, s);                       // greedy.rs:60-123
    // cut rounds (126-190); on large levels, after the first, only the tiles
    // around the last round's moves are recomputed
    const bool inc = g.n >= tune(kTuneIncMinN);
    for (int it = 0; it < niter; ++it)
        refine_pass(g, k, 1, nheavy, heavy_list, B.rb, d_part, s, inc && it > 0, inc && it + 1 < niter);
    if (st) {
        RK_HIP(hipEventRecord(e1, s));
        int32_t passes = 0;
        RK_HIP(hipMemcpyAsync(&passes, B.rb.passes, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        RK_HIP(hipEventSynchronize(e1));
        RK_HIP(hipStreamSynchronize(s));
        RK_HIP(hipEventElapsedTime(&st->refine_ms, e0, e1));
        st->passes = passes;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
}

std::vector<int32_t> grow_seeds(int32_t n, int32_t k, Rng& rng) {
    if (n < k) throw Error(-4, "coarsest graph has fewer vertices than nparts");
    std::vector<int32_t> seeds;
    seeds.reserve(k);
    std::vector<char> used(n, 0);
    while ((int32_t)seeds.size() < k) {                                  // grow.rs:24-32
        const int32_t v = (int32_t)rng.below((uint64_t)n);
        if (!used[v]) {
            used[v] = 1;
            seeds.push_back(v);
        }
    }
    return seeds;
}

void grow_on_device(Context& cx, const std::vector<GrowJob>& jobs) {
    if (jobs.empty()) return;
    // per-job tables one after another; the global scratch only for the
    // graphs too large to grow in LDS
    int64_t sk = 0, skc = 0, sbig = 0, sn = 0;
    size_t lds = 0;
    for (const GrowJob& j : jobs) {
        if (j.k >= (1 << 24)) throw Error(-2, "growing more than 2^24 - 1 parts is not supported");
        const int32_t ncon = j.g.ncon;
        sk += j.k;
        skc += (int64_t)j.k * ncon;
        const size_t need = grow_lds_bytes(j.n, j.m, j.k, ncon);
        if (need <= kGrowLdsBudget) {
            lds = std::max(lds, need);
        } else {
            lds = std::max(lds, grow_global_lds_bytes(j.k));
            sn += j.n;
            if (j.k > 4096) sbig += j.k;
        }
    }
    int32_t* qnext = sn ? cx.arena.get<int32_t>(sn) : nullptr;
    unsigned long long* owner = sn ? cx.arena.get<unsigned long long>(sn) : nullptr;
    int32_t* left = sn ? cx.arena.get<int32_t>(sn) : nullptr;
    int32_t* st = cx.arena.get<int32_t>(7 * sk);
    long long* pw = cx.arena.get<long long>(skc);
    long long* w0 = sbig ? cx.arena.get<long long>(sbig) : nullptr;
    double* tpw = cx.arena.get<double>(skc);
    int32_t* seeds = cx.arena.get<int32_t>(sk);
    double* htpw = (double*)cx.upload_stage(sizeof(double) * (size_t)skc);
    int32_t* hseeds = (int32_t*)cx.upload_stage(sizeof(int32_t) * (size_t)sk);
    GrowArgs* hrows = (GrowArgs*)cx.upload_stage(sizeof(GrowArgs) * jobs.size());
    int64_t ok = 0, okc = 0, obig = 0, on = 0;
    for (size_t i = 0; i < jobs.size(); ++i) {
        const GrowJob& j = jobs[i];
        const int32_t ncon = j.g.ncon;
        std::copy(j.tpw.begin(), j.tpw.end(), htpw + okc);
        std::copy(j.seeds.begin(), j.seeds.end(), hseeds + ok);
        GrowArgs r;
        r.xadj = j.g.xadj;
        r.adjncy = j.g.adjncy;
        r.vwgt = j.g.vwgt;
        r.out = j.out;
        r.v0 = j.v0;
        r.e0 = j.e0;
        r.n = j.n;
        r.m = j.m;
        r.k = j.k;
        r.ncon = ncon;
        r.lds = grow_lds_bytes(j.n, j.m, j.k, ncon) <= kGrowLdsBudget ? 1 : 0;
        r.norep = j.norep ? 1 : 0;
        r.tpw = tpw + okc;
        r.seeds = seeds + ok;
        r.st = st + 7 * ok;
        r.pw = pw + okc;
        if (!r.lds) {
            r.qnext = qnext + on;
            r.owner = owner + on;
            r.leftover = left + on;
            on += j.n;
            if (j.k > 4096) {
                r.w0 = w0 + obig;
                obig += j.k;
            }
        }
        hrows[i] = r;
        ok += j.k;
        okc += (int64_t)j.k * ncon;
    }
    hipStream_t s = cx.stream;
    GrowArgs* d_rows = cx.arena.get<GrowArgs>((int64_t)jobs.size());
    RK_HIP(hipMemcpyAsync(tpw, htpw, sizeof(double) * (size_t)skc, hipMemcpyHostToDevice, s));
    RK_HIP(hipMemcpyAsync(seeds, hseeds, sizeof(int32_t) * (size_t)sk, hipMemcpyHostToDevice, s));
    RK_HIP(hipMemcpyAsync(d_rows, hrows, sizeof(GrowArgs) * jobs.size(), hipMemcpyHostToDevice, s));
    grow_partitions(d_rows, (int32_t)jobs.size(), lds, s);
}

int32_t kway_device(Context& cx, const DevGraph& g, int32_t k, const float* tpwgts, const float* ubvec,
                    const Options& o, int32_t* d_part, KwayStats* st) {
    if (k < 2) throw Error(-2, "nparts must be >= 2");                  // kway.rs:31-33
    if (k > g.n) throw Error(-2, "nparts > nvtxs");                     // kway.rs:34-38
    hipStream_t s = cx.stream;
    const int32_t ncon = g.ncon;
    Buffers B = alloc_buffers(cx, g.n, g.m, k, ncon);
    const bool fm = !o.rtype_greedy;                                    // kway.rs:154-161: Fm refines by k-way FM
    FmBufs F;
    if (fm) F = alloc_fm_buffers(cx, g.n, g.m, k, ncon);
    std::vector<int32_t> maxw0;
    const std::vector<int32_t> total = total_weight_checked(cx, g, B, &maxw0);
    int32_t* fin_heavy = cx.arena.get<int32_t>(g.n);
    const int32_t fin_nheavy = heavy_list_sync(g, fin_heavy, B.rb, s);
    std::vector<double> ub(ncon);
    for (int32_t c = 0; c < ncon; ++c)                                  // kway.rs:44-49
        ub[c] = ubvec ? (double)ubvec[c] : 1.0 + (double)o.ufactor / 1000.0;
    const int32_t limit = std::max(20 * k, 20);                         // coarsen/mod.rs:43
    const size_t kc = (size_t)k * ncon;
    // Matched pairs' weight cap (unit inputs too: their coarse levels are
    // weighted).  Level l's vertices weigh at most 2^l times the heaviest
    // input vertex, so a pair there can reach the cap only once 2^(l+1)
    // times that weight exceeds it: the levels below pass none.
    long long* d_cap = cx.arena.get<long long>(ncon);
    int cap_from = 64;                                                  // first level where the cap can bind
    {
        long long* hcap = (long long*)cx.upload_stage(sizeof(long long) * ncon);
        for (int32_t c = 0; c < ncon; ++c) {
            hcap[c] = match_weight_cap(total[c], g.n, k);
            int l = 0;
            while (l < 62 && ((long long)maxw0[c] << (l + 1)) <= hcap[c]) ++l;
            cap_from = std::min(cap_from, l);
        }
        RK_HIP(hipMemcpyAsync(d_cap, hcap, sizeof(long long) * ncon, hipMemcpyHostToDevice, s));
    }
    std::vector<int32_t> ltot(ncon);
    bool have_best = false;
    int32_t best = 0;
    if (st) *st = KwayStats{};
    ArenaAlloc alloc(cx.arena);
    for (int attempt = 0; attempt < o.ncuts; ++attempt) {
        uint64_t seed = o.seed;
        if (seed == 0) {
            std::random_device rd;
            seed = ((uint64_t)rd() << 32) ^ rd();
        } else {
            seed = o.seed + (uint64_t)attempt;                          // kway.rs:55-61
        }
        Rng rng(seed);
        const Arena::Mark mark = cx.arena.mark();
        struct Level {
            DevGraph g;
            int32_t* cmap;        // fine vertex -> vertex of this level
            int32_t* heavy;       // heavy-vertex list of this level
            int32_t nheavy;
        };
        std::vector<Level> levels;
        std::vector<LevelStat> lst;
        const DevGraph* cur = &g;
        for (;;) {                                                      // coarsen/mod.rs:49-69
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (st) {
                RK_HIP(hipEventCreate(&e0));
                RK_HIP(hipEventCreate(&e1));
                RK_HIP(hipEventRecord(e0, s));
            }
            {
                ProfScope ps(kProfMatch, match_bytes(cur->n, cur->m, cur->adjwgt != nullptr), cur->n, s);
                match_graph(*cur, o.ctype_shem, rng.next_u64(), B.cb, s, nullptr,
                            (int)levels.size() >= cap_from ? d_cap : nullptr);
            }
            int32_t* cmap = cx.arena.get<int32_t>(cur->n);
            int32_t* heavy = cx.arena.get<int32_t>(cur->n);
            DevGraph cg;
            ContractResult cr;
            {
                ProfScope ps(kProfContract, 0.0, cur->n, s);
                cr = contract(*cur, B.cb.match, cmap, cg, false, B.cb, alloc, heavy, s);   // one sync
                ps.bytes = contract_bytes(cur->n, cur->m, ncon, cur->adjwgt != nullptr, cr.cn, cr.cm);
            }
            if (st) {
                LevelStat ls;
                ls.n = cur->n;
                ls.m = cur->m;
                RK_HIP(hipEventRecord(e1, s));
                RK_HIP(hipEventSynchronize(e1));
                RK_HIP(hipEventElapsedTime(&ls.coarsen_ms, e0, e1));
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                lst.push_back(ls);
            }
            const int32_t cn = cr.cn;
            const float ratio = (float)cn / (float)cur->n;              // coarsen/mod.rs:61
            levels.push_back({cg, cmap, heavy, cr.nheavy});
            if (cn <= limit || ratio > 0.95f) break;                    // coarsen/mod.rs:65
            cur = &levels.back().g;
        }
        const DevGraph& cg = levels.back().g;
        if (cg.n < k) throw Error(-4, "coarsest graph has fewer vertices than nparts");
        // Threshold tables of every level (graph/mod.rs:392-410, kway.rs:44-49,
        // 67-69, 86-88) in one upload: level i of the hierarchy (0 = input).
        const size_t nl = levels.size() + 1;
        long long* stage = (long long*)cx.host_stage(2 * kc * nl * sizeof(long long));
        for (size_t i = 0; i < nl; ++i) {
            const DevGraph& lg = i == 0 ? g : levels[i - 1].g;
            for (int c = 0; c < ncon; ++c) ltot[c] = lg.vwgt ? total[c] : lg.n;
            level_thresholds(k, ncon, tpwgts, ltot.data(), ub.data(), stage + 2 * kc * i, stage + 2 * kc * i + kc);
        }
        long long* d_thr = cx.arena.get<long long>((int64_t)(2 * kc * nl));
        RK_HIP(hipMemcpyAsync(d_thr, stage, 2 * kc * nl * sizeof(long long), hipMemcpyHostToDevice, s));
        // initial partition of the coarsest level (kway.rs:66-70): BFS growing
        // on the device from seeds drawn here, or the random parts drawn here
        auto t0 = std::chrono::steady_clock::now();
        int32_t* pc = cx.arena.get<int32_t>(cg.n);
        if (o.iptype == 1) {
            int32_t* hp = (int32_t*)cx.host_stage2(sizeof(int32_t) * cg.n);
            for (int32_t v = 0; v < cg.n; ++v) hp[v] = (int32_t)rng.below((uint64_t)k);   // initial/random.rs:8-10
            RK_HIP(hipMemcpyAsync(pc, hp, sizeof(int32_t) * cg.n, hipMemcpyHostToDevice, s));
        } else {
            for (int c = 0; c < ncon; ++c) ltot[c] = cg.vwgt ? total[c] : cg.n;
            GrowJob job;
            job.g = cg;
            job.n = cg.n;
            job.m = cg.m;
            job.k = k;
            job.out = pc;
            job.tpw.resize(kc);
            for (int32_t p = 0; p < k; ++p)                                 // graph/mod.rs:392-410
                for (int32_t c = 0; c < ncon; ++c)
                    job.tpw[(size_t)p * ncon + c] =
                        (tpwgts ? (double)tpwgts[(size_t)p * ncon + c] : 1.0 / (double)k) * (double)ltot[c];
            job.seeds = grow_seeds(cg.n, k, rng);
            grow_on_device(cx, std::vector<GrowJob>{std::move(job)});
        }
        if (st) RK_HIP(hipStreamSynchronize(s));
        auto t1 = std::chrono::steady_clock::now();
        // refine the coarsest level, then project and refine every finer level
        LevelStat cstat;
        cstat.n = cg.n;
        cstat.m = cg.m;
        // refine_level (kway.rs:145-162): greedy, or k-way FM
        auto refine = [&](const DevGraph& lg, size_t li, int32_t* part, int nh, const int32_t* hl, LevelStat* ls,
                          bool cleared, bool pwv) {
            const long long* thr = d_thr + 2 * kc * li;
            if (!fm) {
                refine_level(cx, lg, k, thr, thr + kc, o.niter, part, nh, hl, B, ls, cleared, pwv);
                return;
            }
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (ls) {
                RK_HIP(hipEventCreate(&e0));
                RK_HIP(hipEventCreate(&e1));
                RK_HIP(hipEventRecord(e0, s));
            }
            fm_refine_level(lg, k, thr, o.niter, part, nh, hl, B.rb, F, s, cleared, pwv);
            if (ls) {
                RK_HIP(hipEventRecord(e1, s));
                RK_HIP(hipEventSynchronize(e1));
                RK_HIP(hipEventElapsedTime(&ls->refine_ms, e0, e1));
                ls->passes = o.niter;
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
            }
        };
        refine(cg, nl - 1, pc, levels.back().nheavy, levels.back().heavy, st ? &cstat : nullptr, false, false);
        // With explicit vertex weights a coarse vertex weighs the sum of its
        // fine vertices (graph/mod.rs:284-384), so the projected partition has
        // exactly the part weights the coarser level's last pass left in
        // B.rb.pw (int32 sums wrap alike in any order): the table is carried
        // over instead of recomputed.  Unit weights (vwgt absent) count
        // vertices, which differ from level to level.
        const bool keep_pw = g.vwgt != nullptr && !(debug_flags() & kDebugNoPwCarry);
        for (int64_t i = (int64_t)levels.size() - 1; i >= 0; --i) {     // kway.rs:76-92
            const DevGraph& fine = i == 0 ? g : levels[i - 1].g;
            const int32_t fheavy = i == 0 ? fin_nheavy : levels[i - 1].nheavy;
            const int32_t* flist = i == 0 ? fin_heavy : levels[i - 1].heavy;
            int32_t* pf = cx.arena.get<int32_t>(fine.n);
            project(fine.n, levels[i].cmap, pc, pf, s, keep_pw ? nullptr : B.rb.pw, kc, B.rb.done, 1);
            refine(fine, (size_t)i, pf, fheavy, flist, st ? &lst[i] : nullptr, true, keep_pw);
            pc = pf;
        }
        edge_cut(g, pc, B.cut_acc, B.cut_out, pick_group(g.n, g.m), s);    // kway.rs:94
        RK_HIP(hipMemcpyAsync(B.rb.host_small, B.cut_out, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        if (fm) RK_HIP(hipMemcpyAsync(B.rb.host_small + 1, F.err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        RK_HIP(hipStreamSynchronize(s));
        const int32_t obj = B.rb.host_small[0];
        if (fm && B.rb.host_small[1]) throw Error(-4, "FM pass exceeded its pop budget");
        if (!have_best || obj < best) {                                 // kway.rs:96-98
            have_best = true;
            best = obj;
            RK_HIP(hipMemcpyAsync(d_part, pc, sizeof(int32_t) * g.n, hipMemcpyDeviceToDevice, s));
        }
        if (st) {
            lst.push_back(cstat);
            st->levels = lst;
            st->coarsest_n = cg.n;
            st->initial_ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
            st->attempts = attempt + 1;
        }
        cx.arena.rewind(mark);
    }
    RK_HIP(hipStreamSynchronize(s));
    return best;
}

}  // namespace rk
