// batch.cpp -- many k-way partitions in one pass over a segmented union graph.
//
// Every graph of the batch is partitioned exactly as kway_device partitions
// it alone (partition/kway.rs:20-141): the same matching keys and seeds
// (graph-local ids, the graph's own seed stream), the same coarsening stop
// rule (coarsen/mod.rs:49-69), the same growing of its coarsest graph
// (initial/grow.rs:12-103), the same greedy passes (greedy.rs:36-193) and
// the same best-of-ncuts choice (kway.rs:96-98).  What changes is the
// schedule: the attempts of all graphs run level by level in lockstep (as
// many attempts per pass as the int32 offsets and device memory allow, each
// on its own copy of the batch), so each kernel launch covers every
// (graph, attempt) still present at that level.
//
// Level l holds the graphs that reached it laid out one after another.  A
// graph whose coarsening stops at level l is present at l but not matched
// there (its vertices get cmap -1 and leave the hierarchy); its grown
// partition is scattered into level l's part array before refinement, while
// the other graphs of level l receive theirs by projection from level l+1.
// Tiles of 64 vertices never straddle graphs and start at a multiple of 32
// per graph, so every graph owns whole words of the tile bitmap and is
// admitted by its own workgroup.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <random>
#include <thread>

#include "driver.hpp"
#include "profiler.hpp"

namespace rk {

namespace {

// Asynchronous upload of v through the pinned bump area (into d, or into a
// new arena array).
template <typename T>
T* put(Context& cx, const std::vector<T>& v, T* d = nullptr) {
    if (!d) d = cx.arena.get<T>((int64_t)v.size());
    if (!v.empty()) {
        T* h = (T*)cx.upload_stage(sizeof(T) * v.size());
        std::copy(v.begin(), v.end(), h);
        RK_HIP(hipMemcpyAsync(d, h, sizeof(T) * v.size(), hipMemcpyHostToDevice, cx.stream));
    }
    return d;
}

// Per-graph extent of one level (zero n for graphs not present).
struct Extent {
    int32_t v0 = 0, n = 0, e0 = 0, m = 0;
};

struct BLevel {
    DevGraph g;
    int32_t* gseg = nullptr;      // [g.n] graph of each vertex
    int32_t* cmap = nullptr;      // previous level's vertex -> this level (null at level 0)
    int32_t* heavy = nullptr;
    int32_t nheavy = 0;
    int32_t* d_gv0 = nullptr;     // [G] first vertex of each graph here
    std::vector<Extent> ext;
    std::vector<uint8_t> on;      // matched and contracted at this level
    std::vector<uint8_t> last;    // this level is the graph's coarsest
};

std::vector<Extent> read_extents(Context& cx, const DevGraph& g, const int32_t* gseg, int32_t G, int32_t* d_tmp) {
    graph_ranges(g.n, gseg, g.xadj, G, d_tmp, cx.stream);
    int32_t* h = (int32_t*)cx.host_stage2(sizeof(int32_t) * 4 * (size_t)G);
    RK_HIP(hipMemcpyAsync(h, d_tmp, sizeof(int32_t) * 4 * (size_t)G, hipMemcpyDeviceToHost, cx.stream));
    RK_HIP(hipStreamSynchronize(cx.stream));
    std::vector<Extent> e(G);
    for (int32_t i = 0; i < G; ++i) {
        e[i].v0 = h[4 * i];
        e[i].n = h[4 * i + 1] - h[4 * i];
        e[i].e0 = h[4 * i + 2];
        e[i].m = h[4 * i + 3] - h[4 * i + 2];
    }
    return e;
}

// The 2-way FM tables of a level: the graphs present, each with its tree
// region, and per tree level the prefix of their node counts.
Fm2Level fm2_level(Context& cx, const int32_t* gseg, const std::vector<Extent>& ext, int32_t G) {
    std::vector<Fm2Graph> gd;
    std::vector<int32_t> of(G, -1);
    int64_t tb = 0;
    int nlvl = 1;
    int32_t max_n = 0;
    for (int32_t i = 0; i < G; ++i) {
        if (ext[i].n <= 0) continue;
        Fm2Graph d;
        d.g = i;
        d.v0 = ext[i].v0;
        d.n = ext[i].n;
        d.tbase = tb;
        of[i] = (int32_t)gd.size();
        gd.push_back(d);
        tb += fm2_tree_nodes(ext[i].n);
        nlvl = std::max(nlvl, (int)fm2_tree_levels(ext[i].n));
        max_n = std::max(max_n, ext[i].n);
    }
    Fm2Level lv;
    lv.max_n = max_n;
    lv.ngraphs = (int32_t)gd.size();
    lv.gtotal = G;
    lv.nlvl = nlvl;
    lv.gseg = gseg;
    const size_t ng = gd.size();
    std::vector<int64_t> nodes((size_t)kFm2MaxLvl * (ng + 1), 0);
    for (int l = 1; l < nlvl; ++l) {
        int64_t acc = 0;
        for (size_t i = 0; i < ng; ++i) {
            nodes[(size_t)l * (ng + 1) + i] = acc;
            acc += fm2_level_nodes(gd[i].n, l);   // 0 when the graph's tree is shorter
        }
        nodes[(size_t)l * (ng + 1) + ng] = acc;
        lv.level_nodes[l] = acc;
    }
    lv.graphs = put(cx, gd);
    lv.gdesc_of = put(cx, of);
    lv.nodes = put(cx, nodes);
    return lv;
}

inline int64_t padded_tiles(int32_t n) { return ((((int64_t)n + kTile - 1) / kTile) + 31) & ~(int64_t)31; }

// One attempt of every graph of the union g: graph i runs with the seed
// seeds[i] as given (the caller derives it per attempt) and, when tpw is
// given, the target fractions tpw[gkc[i] + p*ncon + c] of its parts (the
// k*ncon tables of the graphs one after another; else 1/k each).  Greedy
// refinement, or 2-way FM (refine/fm.rs) when every graph has two parts.
// Writes d_part[g.n] and objval[G] (host).
void batch_attempt(Context& cx, const DevGraph& g, int32_t G, const int32_t* gvtx, const int32_t* nparts,
                   const float* ubvec, const Options& o, const uint64_t* seeds, int32_t* d_part, int32_t* objval,
                   KwayStats* st, const float* tpw = nullptr) {
    const int32_t ncon = g.ncon;
    std::vector<int32_t> gk(G);
    std::vector<int32_t> gkc(G);
    int64_t kc_total = 0;
    int32_t kmax = 0;
    int64_t tiles0 = 0;
    for (int32_t i = 0; i < G; ++i) {
        const int32_t n = gvtx[i + 1] - gvtx[i];
        if (n <= 0) throw Error(-2, "graph " + std::to_string(i) + " has no vertices");
        if (nparts[i] < 2) throw Error(-2, "graph " + std::to_string(i) + ": nparts must be >= 2");   // kway.rs:31-33
        if (nparts[i] > n) throw Error(-2, "graph " + std::to_string(i) + ": nparts > nvtxs");        // kway.rs:34-38
        gk[i] = nparts[i];
        gkc[i] = (int32_t)kc_total;
        kc_total += (int64_t)nparts[i] * ncon;
        if (kc_total > INT32_MAX) throw Error(-2, "sum of nparts*ncon over the batch exceeds the int32 range");
        kmax = std::max(kmax, nparts[i]);
        tiles0 += padded_tiles(n);
    }
    if (tiles0 > INT32_MAX / kTile) throw Error(-2, "batch too large for the tile tables");
    if (!o.rtype_greedy && kmax != 2) throw Error(-2, "batched FM refinement is the 2-way FM of bisections");
    hipStream_t s = cx.stream;
    ArenaAlloc alloc(cx.arena);
    Buffers B = alloc_buffers(cx, g.n, g.m, kmax, ncon, tiles0, kc_total);
    // per-graph device tables that live for the whole call
    int32_t* d_gk = put(cx, gk);
    int32_t* d_gkc = put(cx, gkc);
    int32_t* d_gvtx = put(cx, std::vector<int32_t>(gvtx, gvtx + G + 1));
    int32_t* d_gdone = cx.arena.get<int32_t>(4 * (int64_t)G);
    int32_t* d_gnmoves = d_gdone + G;
    int32_t* d_gpasses = d_gdone + 2 * (int64_t)G;
    int32_t* d_gfull = d_gdone + 3 * (int64_t)G;
    RK_HIP(hipMemsetAsync(d_gdone, 0, sizeof(int32_t) * 4 * (size_t)G, s));
    int32_t* d_ranges = cx.arena.get<int32_t>(4 * (int64_t)G);
    unsigned long long* d_cutacc = cx.arena.get<unsigned long long>(G);
    int32_t* d_cut = cx.arena.get<int32_t>(G);
    int32_t* gseg0 = cx.arena.get<int32_t>(g.n);
    {
        std::vector<int32_t> gn(G);
        for (int32_t i = 0; i < G; ++i) gn[i] = gvtx[i + 1] - gvtx[i];
        fill_gseg(G, d_gvtx, put(cx, gn), gseg0, s);
    }
    // per-graph totals (graph/mod.rs:392-410).  With explicit vertex weights
    // they, and so the thresholds, are the same at every level; without
    // (every coarse vertex then counts 1, as the reference's None weights
    // stay None when contracted) a level's total is its vertex count and
    // the thresholds are rebuilt per level.
    const bool unit = g.vwgt == nullptr;
    std::vector<std::vector<int32_t>> total(G, std::vector<int32_t>(ncon));
    std::vector<long long> gmaxw;                 // heaviest input vertex per graph and constraint
    if (unit) {
        for (int32_t i = 0; i < G; ++i)
            for (int c = 0; c < ncon; ++c) total[i][c] = gvtx[i + 1] - gvtx[i];
    } else {
        long long* d_sum = cx.arena.get<long long>(2 * (int64_t)G * ncon);
        graph_column_sums(g, d_gvtx, G, d_sum, s, d_sum + (int64_t)G * ncon);
        long long* h = (long long*)cx.host_stage2(sizeof(long long) * 2 * (size_t)G * ncon);
        RK_HIP(hipMemcpyAsync(h, d_sum, sizeof(long long) * 2 * (size_t)G * ncon, hipMemcpyDeviceToHost, s));
        RK_HIP(hipStreamSynchronize(s));
        gmaxw.assign(h + (size_t)G * ncon, h + 2 * (size_t)G * ncon);
        for (int32_t i = 0; i < G; ++i)
            for (int c = 0; c < ncon; ++c) {
                const long long v = h[(size_t)i * ncon + c];
                if (v > INT32_MAX)
                    throw Error(-4, "graph " + std::to_string(i) +
                                        ": total vertex weight of a constraint exceeds the int32 range");
                total[i][c] = (int32_t)v;
            }
    }
    std::vector<double> ub(ncon);
    for (int32_t c = 0; c < ncon; ++c) ub[c] = ubvec ? (double)ubvec[c] : 1.0 + (double)o.ufactor / 1000.0;
    // floor(target*ub) tables of the graphs (kway.rs:44-49, 67-69, 86-88);
    // ltot(i) gives graph i's per-constraint totals at the level
    auto upload_thresholds = [&](auto ltot) {
        std::vector<long long> fm(kc_total), om(kc_total);
        std::vector<int32_t> t(ncon);
        for (int32_t i = 0; i < G; ++i) {
            ltot(i, t.data());
            level_thresholds(gk[i], ncon, tpw ? tpw + gkc[i] : nullptr, t.data(), ub.data(), fm.data() + gkc[i],
                             om.data() + gkc[i]);
        }
        put(cx, fm, B.rb.fmax);
        put(cx, om, B.rb.omax);
    };
    if (!unit) upload_thresholds([&](int32_t i, int32_t* t) { std::copy(total[i].begin(), total[i].end(), t); });
    // The matched pairs' weight cap of every graph (match_weight_cap) and the
    // first level where it can bind (a level-l vertex weighs at most 2^l
    // times the graph's heaviest input vertex); a level's table holds the
    // caps of the graphs where they can bind, INT64_MAX for the others.
    if (unit) gmaxw.assign((size_t)G * ncon, 1);
    std::vector<long long> gcap((size_t)G * ncon);
    std::vector<int> gcap_from(G, 64);
    for (int32_t i = 0; i < G; ++i)
        for (int c = 0; c < ncon; ++c) {
            const long long cp = match_weight_cap(total[i][c], gvtx[i + 1] - gvtx[i], gk[i]);
            gcap[(size_t)i * ncon + c] = cp;
            int l = 0;
            while (l < 62 && (gmaxw[(size_t)i * ncon + c] << (l + 1)) <= cp) ++l;
            gcap_from[i] = std::min(gcap_from[i], l);
        }
    auto level_caps = [&](size_t l) -> long long* {
        std::vector<long long> t((size_t)G * ncon, LLONG_MAX);
        bool any = false;
        for (int32_t i = 0; i < G; ++i)
            if ((int)l >= gcap_from[i]) {
                any = true;
                for (int c = 0; c < ncon; ++c) t[(size_t)i * ncon + c] = gcap[(size_t)i * ncon + c];
            }
        return any ? put(cx, t) : nullptr;
    };
    // 2-way FM work buffers (trees sized by the finest level's graphs)
    Fm2Bufs F2{};
    if (!o.rtype_greedy) {
        int64_t tree = 0;
        for (int32_t i = 0; i < G; ++i) tree += fm2_tree_nodes(gvtx[i + 1] - gvtx[i]);
        F2.leaf = cx.arena.get<uint64_t>(g.n);
        F2.tree = cx.arena.get<uint64_t>(tree);
        F2.state = (uint8_t*)cx.arena.get<int32_t>(((int64_t)g.n + 3) / 4);
        F2.gain = cx.arena.get<int32_t>(g.n);
        F2.ext = cx.arena.get<int32_t>(g.n);
        F2.moves = cx.arena.get<int32_t>(g.n);
        int32_t* symf = cx.arena.get<int32_t>(1);
        fm2_symmetry(g, symf, s);             // holds for every level: contraction keeps it
        F2.sym = symf;
        int32_t* pg = cx.arena.get<int32_t>(5 * (int64_t)G + 1);
        F2.cut0 = pg;
        F2.cut1 = pg + G;
        F2.best_len = pg + 2 * (int64_t)G;
        F2.best_cut = pg + 3 * (int64_t)G;
        F2.nmoves = pg + 4 * (int64_t)G;
        F2.err = pg + 5 * (int64_t)G;
        RK_HIP(hipMemsetAsync(F2.err, 0, sizeof(int32_t), s));
    }
    int32_t* fin_heavy = cx.arena.get<int32_t>(g.n);
    const int32_t fin_nheavy = heavy_list_sync(g, fin_heavy, B.rb, s);
    const std::vector<Extent> ext0 = read_extents(cx, g, gseg0, G, d_ranges);
    int32_t* d_gv00 = nullptr;
    {
        std::vector<int32_t> v0(G);
        for (int32_t i = 0; i < G; ++i) v0[i] = gvtx[i];
        d_gv00 = put(cx, v0);
    }
    if (st) *st = KwayStats{};
    {
        std::vector<Rng> rng;
        rng.reserve(G);
        for (int32_t i = 0; i < G; ++i) rng.emplace_back(seeds[i]);
        std::vector<BLevel> lv(1);
        lv[0].g = g;
        lv[0].gseg = gseg0;
        lv[0].heavy = fin_heavy;
        lv[0].nheavy = fin_nheavy;
        lv[0].d_gv0 = d_gv00;
        lv[0].ext = ext0;
        lv[0].on.assign(G, 1);
        lv[0].last.assign(G, 0);
        std::vector<LevelStat> lst;
        // ---- coarsening (coarsen/mod.rs:49-69 per graph)
        for (size_t l = 0;; ++l) {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (st) {
                RK_HIP(hipEventCreate(&e0));
                RK_HIP(hipEventCreate(&e1));
                RK_HIP(hipEventRecord(e0, s));
            }
            BLevel& L = lv[l];
            cx.set_phase(kPhaseCoarsen, (int32_t)l);
            std::vector<uint64_t> gseed(G, 0);
            for (int32_t i = 0; i < G; ++i)
                if (L.on[i]) gseed[i] = rng[i].next_u64();
            GraphSegs gs;
            gs.gseg = L.gseg;
            gs.gv0 = L.d_gv0;
            gs.gseed = put(cx, gseed);
            gs.gon = put(cx, L.on);
            gs.cgseg = cx.arena.get<int32_t>(L.g.n);
            {
                ProfScope ps(kProfMatch, match_bytes(L.g.n, L.g.m, L.g.adjwgt != nullptr), L.g.n, s);
                match_graph(L.g, o.ctype_shem, 0, B.cb, s, &gs, level_caps(l));
            }
            BLevel C;
            C.cmap = cx.arena.get<int32_t>(L.g.n);
            C.heavy = cx.arena.get<int32_t>(L.g.n);
            ContractResult cr;
            {
                ProfScope ps(kProfContract, 0.0, L.g.n, s);
                cr = contract(L.g, B.cb.match, C.cmap, C.g, false, B.cb, alloc, C.heavy, s, &gs);
                ps.bytes = contract_bytes(L.g.n, L.g.m, ncon, L.g.adjwgt != nullptr, cr.cn, cr.cm);
            }
            C.nheavy = cr.nheavy;
            C.gseg = gs.cgseg;
            C.ext = read_extents(cx, C.g, C.gseg, G, d_ranges);
            C.on.assign(G, 0);
            C.last.assign(G, 0);
            bool any = false;
            std::vector<int32_t> v0(G, 0);
            for (int32_t i = 0; i < G; ++i) {
                v0[i] = C.ext[i].v0;
                if (!L.on[i]) continue;
                const int32_t cn = C.ext[i].n;
                const int32_t limit = std::max(20 * gk[i], 20);                 // coarsen/mod.rs:43
                const float ratio = (float)cn / (float)L.ext[i].n;              // coarsen/mod.rs:61
                if (cn <= limit || ratio > 0.95f) {                              // coarsen/mod.rs:65
                    if (cn < gk[i])
                        throw Error(-4, "graph " + std::to_string(i) +
                                            ": coarsest graph has fewer vertices than nparts");
                    C.last[i] = 1;
                } else {
                    C.on[i] = 1;
                    any = true;
                }
            }
            C.d_gv0 = put(cx, v0);
            if (st) {
                LevelStat ls;
                ls.n = L.g.n;
                ls.m = L.g.m;
                RK_HIP(hipEventRecord(e1, s));
                RK_HIP(hipEventSynchronize(e1));
                RK_HIP(hipEventElapsedTime(&ls.coarsen_ms, e0, e1));
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                lst.push_back(ls);
            }
            lv.push_back(std::move(C));
            if (!any) break;
        }
        const size_t nl = lv.size();
        // ---- initial partitions of the coarsest graphs (kway.rs:66-70): BFS
        // growing on the device from seeds drawn here per graph, or the
        // random parts drawn here (uploaded in one copy)
        cx.set_phase(kPhaseInitial);
        auto t0 = std::chrono::steady_clock::now();
        std::vector<int32_t> lvl_of(G, -1);
        std::vector<int64_t> part_off(G, 0);
        int64_t part_tot = 0;
        std::vector<GrowJob> jobs;
        for (size_t l = 1; l < nl; ++l)
            for (int32_t i = 0; i < G; ++i) {
                if (!lv[l].last[i]) continue;
                const Extent& e = lv[l].ext[i];
                lvl_of[i] = (int32_t)l;
                part_off[i] = part_tot;
                part_tot += e.n;
                if (e.n < gk[i])
                    throw Error(-4, "graph " + std::to_string(i) + ": coarsest graph has fewer vertices than nparts");
                if (o.iptype == 1) continue;
                GrowJob j;
                j.g = lv[l].g;
                j.v0 = e.v0;
                j.e0 = e.e0;
                j.n = e.n;
                j.m = e.m;
                j.k = gk[i];
                j.tpw.resize((size_t)gk[i] * ncon);
                for (int32_t p = 0; p < gk[i]; ++p)                             // graph/mod.rs:392-410
                    for (int32_t c = 0; c < ncon; ++c) {
                        const double f = tpw ? (double)tpw[gkc[i] + p * ncon + c] : 1.0 / (double)gk[i];
                        j.tpw[(size_t)p * ncon + c] = f * (double)(unit ? e.n : total[i][c]);
                    }
                j.seeds = grow_seeds(e.n, gk[i], rng[i]);                       // grow.rs:24-32
                jobs.push_back(std::move(j));
            }
        // the coarsest graphs' parts, graph after graph (grown in one launch
        // over every level, or drawn here), scattered into each level's part
        // array before its refinement
        int32_t* d_gparts = cx.arena.get<int32_t>(std::max<int64_t>(part_tot, 1));
        if (o.iptype == 1) {
            int32_t* hpart = (int32_t*)cx.upload_stage(sizeof(int32_t) * (size_t)std::max<int64_t>(part_tot, 1));
            for (int32_t i = 0; i < G; ++i) {
                if (lvl_of[i] < 0) continue;
                int32_t* hp = hpart + part_off[i];
                for (int32_t v = 0; v < lv[lvl_of[i]].ext[i].n; ++v)
                    hp[v] = (int32_t)rng[i].below((uint64_t)gk[i]);           // initial/random.rs:8-10
            }
            RK_HIP(hipMemcpyAsync(d_gparts, hpart, sizeof(int32_t) * (size_t)part_tot, hipMemcpyHostToDevice, s));
        } else {
            size_t jj = 0;
            for (size_t l = 1; l < nl; ++l)
                for (int32_t i = 0; i < G; ++i)
                    if (lv[l].last[i]) jobs[jj++].out = d_gparts + part_off[i];
            grow_on_device(cx, jobs);
        }
        if (st) RK_HIP(hipStreamSynchronize(s));
        auto t1 = std::chrono::steady_clock::now();
        // ---- refinement from the deepest level up (kway.rs:73-92)
        int32_t* pc = nullptr;
        std::vector<LevelStat> rst(nl);
        for (int64_t l = (int64_t)nl - 1; l >= 0; --l) {
            BLevel& L = lv[l];
            cx.set_phase(kPhaseRefine, (int32_t)l);
            int32_t* pf = cx.arena.get<int32_t>(L.g.n);
            if (l + 1 < (int64_t)nl) project(L.g.n, lv[l + 1].cmap, pc, pf, s);
            std::vector<int32_t> goff(G, -1);
            bool anylast = false;
            for (int32_t i = 0; i < G; ++i)
                if (lvl_of[i] == l) {
                    goff[i] = (int32_t)part_off[i];
                    anylast = true;
                }
            if (anylast) scatter_parts(L.g.n, L.gseg, L.d_gv0, put(cx, goff), d_gparts, pf, s);
            // tile and admission tables of the graphs present
            std::vector<int4> tdesc;
            std::vector<AdmitArgs> adm;
            int64_t tb = 0;
            for (int32_t i = 0; i < G; ++i) {
                const Extent& e = L.ext[i];
                if (e.n == 0) continue;
                tdesc.push_back(make_int4(i, e.v0, e.n, (int32_t)tb));
                AdmitArgs aa{};
                aa.k = gk[i];
                aa.ncon = ncon;
                aa.ntiles = (int32_t)padded_tiles(e.n);
                aa.tile_cnt = B.rb.tile_cnt + tb;
                aa.tile_bits = B.rb.tile_bits + tb / 32;
                aa.tile_list = B.rb.tile_off + tb;
                aa.tile_rec = B.rb.tile_rec + tb * kTile;
                aa.mov = B.rb.mov + e.v0;
                aa.pool = B.rb.pool;
                aa.xadj = L.g.xadj;
                aa.vwgt = L.g.vwgt;
                aa.fmax = B.rb.fmax + gkc[i];
                aa.omax = B.rb.omax + gkc[i];
                aa.pw = B.rb.pw + gkc[i];
                aa.part = pf;
                aa.moves = B.rb.moves + e.v0;
                aa.nmoves = d_gnmoves + i;
                aa.done = d_gdone + i;
                aa.passes = d_gpasses + i;
                aa.gscratch = B.rb.adm_scratch + 3 * (int64_t)gkc[i];
                aa.adjncy = L.g.adjncy;
                aa.tile_dirty = B.rb.tile_dirty + tb / 32;
                aa.full = d_gfull + i;
                aa.v0 = e.v0;
                adm.push_back(aa);
                tb += padded_tiles(e.n);
            }
            BatchLevel bl;
            bl.ntiles = (int32_t)tb;
            int32_t* tt = cx.arena.get<int32_t>(3 * tb);
            bl.tile_v0 = tt;
            bl.tile_nv = tt + tb;
            bl.tile_g = tt + 2 * tb;
            fill_tiles(put(cx, tdesc), (int32_t)tdesc.size(), tt, tt + tb, tt + 2 * tb, s);
            bl.gseg = L.gseg;
            bl.gkc = d_gkc;
            bl.gk = d_gk;
            bl.gdone = d_gdone;
            bl.gfull = d_gfull;
            bl.admit = put(cx, adm);
            bl.ngraphs = (int32_t)adm.size();
            bl.kmax = kmax;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (st) {
                RK_HIP(hipEventCreate(&e0));
                RK_HIP(hipEventCreate(&e1));
                RK_HIP(hipEventRecord(e0, s));
            }
            if (unit)
                upload_thresholds([&](int32_t i, int32_t* t) {
                    for (int c = 0; c < ncon; ++c) t[c] = L.ext[i].n > 0 ? L.ext[i].n : 1;
                });
            // greedy.rs:50.  With explicit vertex weights a graph projected from
            // the coarser level keeps that level's final table (coarse weights
            // are sums of fine ones, so the projected partition's table is
            // the same); only the graphs grown at this level are summed.
            if (o.rtype_greedy && !unit && l + 1 < (int64_t)nl) {
                std::vector<uint8_t> mask(G, 0);
                for (int32_t i = 0; i < G; ++i) mask[i] = lvl_of[i] == l ? 1 : 0;
                part_weights_batch(L.g, pf, L.gseg, d_gkc, d_gk, kc_total, B.rb.pw, s, put(cx, mask), G);
            } else {
                part_weights_batch(L.g, pf, L.gseg, d_gkc, d_gk, kc_total, B.rb.pw, s);
            }
            RK_HIP(hipMemsetAsync(d_gdone, 0, sizeof(int32_t) * (size_t)G, s));
            if (o.rtype_greedy) {
                refine_pass_batch(L.g, bl, 0, L.nheavy, L.heavy, B.rb, pf, s);              // greedy.rs:60-123
                for (int it = 0; it < o.niter; ++it)                                       // greedy.rs:126-190
                    refine_pass_batch(L.g, bl, 1, L.nheavy, L.heavy, B.rb, pf, s, it > 0);
            } else {                                                                       // refine/mod.rs:179-182
                const Fm2Level fl = fm2_level(cx, L.gseg, L.ext, G);
                fm2_refine_batch(L.g, fl, F2, d_gkc, B.rb.fmax, B.rb.pw, pf, d_gdone, o.niter, d_cutacc, s);
            }
            if (st) {
                RK_HIP(hipEventRecord(e1, s));
                RK_HIP(hipEventSynchronize(e1));
                rst[l].n = L.g.n;
                rst[l].m = L.g.m;
                RK_HIP(hipEventElapsedTime(&rst[l].refine_ms, e0, e1));
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
            }
            pc = pf;
        }
        // ---- per-graph cut (kway.rs:94)
        cx.set_phase(kPhaseFinish);
        edge_cut_batch(g, pc, gseg0, G, d_cutacc, d_cut, s);
        int32_t* hc = (int32_t*)cx.host_stage2(sizeof(int32_t) * (size_t)G);
        RK_HIP(hipMemcpyAsync(hc, d_cut, sizeof(int32_t) * (size_t)G, hipMemcpyDeviceToHost, s));
        RK_HIP(hipMemcpyAsync(d_part, pc, sizeof(int32_t) * (size_t)g.n, hipMemcpyDeviceToDevice, s));
        RK_HIP(hipStreamSynchronize(s));
        for (int32_t i = 0; i < G; ++i) objval[i] = hc[i];
        if (st) {
            st->levels.clear();
            for (size_t l = 0; l < nl; ++l) {
                LevelStat ls = rst[l];
                if (l < lst.size()) ls.coarsen_ms = lst[l].coarsen_ms;
                ls.n = lv[l].g.n;
                ls.m = lv[l].g.m;
                st->levels.push_back(ls);
            }
            st->coarsest_n = (int32_t)part_tot;
            st->initial_ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
            st->attempts = 1;
        }
    }
}

}  // namespace

void fm2_refine_device(Context& cx, const DevGraph& g, const long long* d_fmax, int niter, int32_t* part) {
    hipStream_t s = cx.stream;
    int32_t* gseg = cx.arena.get<int32_t>(g.n);
    RK_HIP(hipMemsetAsync(gseg, 0, sizeof(int32_t) * (size_t)g.n, s));
    std::vector<Extent> ext(1);
    ext[0].n = g.n;
    const Fm2Level fl = fm2_level(cx, gseg, ext, 1);
    Fm2Bufs F2{};
    F2.leaf = cx.arena.get<uint64_t>(g.n);
    F2.tree = cx.arena.get<uint64_t>(fm2_tree_nodes(g.n));
    F2.state = (uint8_t*)cx.arena.get<int32_t>(((int64_t)g.n + 3) / 4);
    F2.gain = cx.arena.get<int32_t>(g.n);
    F2.ext = cx.arena.get<int32_t>(g.n);
    F2.moves = cx.arena.get<int32_t>(g.n);
    int32_t* symf = cx.arena.get<int32_t>(1);
    fm2_symmetry(g, symf, s);
    F2.sym = symf;
    int32_t* pg = cx.arena.get<int32_t>(16);
    F2.cut0 = pg;
    F2.cut1 = pg + 1;
    F2.best_len = pg + 2;
    F2.best_cut = pg + 3;
    F2.nmoves = pg + 4;
    F2.err = pg + 5;
    int32_t* gdone = pg + 6;
    int32_t* gkc = pg + 7;
    int32_t* gk = pg + 8;
    RK_HIP(hipMemsetAsync(pg, 0, sizeof(int32_t) * 16, s));
    put(cx, std::vector<int32_t>{2}, gk);
    int32_t* pw = cx.arena.get<int32_t>(2 * (int64_t)g.ncon);
    unsigned long long* acc = cx.arena.get<unsigned long long>(1);
    part_weights_batch(g, part, gseg, gkc, gk, 2 * (int64_t)g.ncon, pw, s);   // fm.rs:20
    fm2_refine_batch(g, fl, F2, gkc, d_fmax, pw, part, gdone, niter, acc, s);
    int32_t err = 0;
    RK_HIP(hipMemcpyAsync(&err, F2.err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RK_HIP(hipStreamSynchronize(s));
    if (err) throw Error(-4, "2-way FM pass exceeded its pop budget");
}

void bisect_batch_device(Context& cx, const DevGraph& g, int32_t G, const int32_t* gvtx, const float* tpw2,
                         const float* ubvec, const Options& o, const uint64_t* seeds, int32_t* d_part) {
    if (G <= 0) return;
    if (gvtx[0] != 0 || gvtx[G] != g.n) throw Error(-4, "bisection batch offsets do not cover the union");
    const std::vector<int32_t> two(G, 2);
    std::vector<int32_t> cut(G);
    batch_attempt(cx, g, G, gvtx, two.data(), ubvec, o, seeds, d_part, cut.data(), nullptr, tpw2);
}

int32_t batch_attempts_per_call(int64_t n, int64_t m, int64_t tiles, int64_t kc_total, int32_t ngraphs, int ncuts,
                                int max_pack, size_t reusable) {
    int64_t p = ncuts;
    if (max_pack > 0) p = std::min<int64_t>(p, max_pack);
    // int32 vertex ids, CSR offsets, tile records and threshold offsets
    p = std::min<int64_t>(p, (INT32_MAX - 1) / std::max<int64_t>(n, 1));
    p = std::min<int64_t>(p, INT32_MAX / std::max<int64_t>(m, 1));
    p = std::min<int64_t>(p, (INT32_MAX / kTile) / std::max<int64_t>(tiles, 1));
    p = std::min<int64_t>(p, INT32_MAX / std::max<int64_t>(kc_total, 1));
    p = std::min<int64_t>(p, INT32_MAX / std::max<int64_t>(ngraphs, 1));
    // device memory per copy: the work buffers (36 B per entry: candidate
    // pool, contraction staging and hash tables; about 120 B per vertex) and
    // the hierarchy (the copy itself, the coarse levels, per-level maps and
    // parts), about 56 B per entry and 200 B per vertex in all, within 60 %
    // of the device's free memory plus what the arena already holds
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) freeb = 0;
    const int64_t per_copy = 200 * n + 56 * m + (1 << 20);
    const int64_t room = (int64_t)((double)(freeb + reusable) * 0.6);
    p = std::min<int64_t>(p, std::max<int64_t>(room / per_copy, 1));
    p = std::max<int64_t>(p, 1);
    // balanced groups: the fewest calls, then the same attempts in each
    const int64_t groups = (ncuts + p - 1) / p;
    return (int32_t)((ncuts + groups - 1) / groups);
}

void kway_batch_device(Context& cx, const DevGraph& g, int32_t G, const int32_t* gvtx, const int32_t* nparts,
                       const float* ubvec, const Options& o, const uint64_t* seeds, int32_t* d_part, int32_t* objval,
                       KwayStats* st, int max_pack) {
    if (G <= 0) throw Error(-2, "a batch needs at least one graph");
    if (!o.rtype_greedy)
        throw Error(-2, "the batched and served forms refine greedily: set options[METIS_OPTION_RTYPE]=1 "
                        "(Greedy) and options[METIS_OPTION_PTYPE]=0 so that the array is read; k-way FM runs "
                        "through RMETIS_PartGraphKway and rmetis_hip_kway_device");
    if (gvtx[0] != 0 || gvtx[G] != g.n) throw Error(-2, "graph offsets must start at 0 and end at nvtxs");
    int64_t tiles0 = 0, kc_total = 0;
    for (int32_t i = 0; i < G; ++i) {
        const int32_t n = gvtx[i + 1] - gvtx[i];
        if (n <= 0) throw Error(-2, "graph " + std::to_string(i) + " has no vertices");
        if (nparts[i] < 2) throw Error(-2, "graph " + std::to_string(i) + ": nparts must be >= 2");   // kway.rs:31-33
        if (nparts[i] > n) throw Error(-2, "graph " + std::to_string(i) + ": nparts > nvtxs");        // kway.rs:34-38
        tiles0 += padded_tiles(n);
        kc_total += (int64_t)nparts[i] * g.ncon;
    }
    // Attempt a of graph i runs with seed + a (kway.rs:55-61).  The attempts
    // are independent, so P of them run in one pass over a union holding P
    // copies of the batch (copy j is attempt a0 + j of every graph): every
    // level kernel then covers P times the graphs.  Each graph keeps its
    // strictly best attempt, the earliest among equal cuts (kway.rs:96-98).
    const int32_t P =
        batch_attempts_per_call(g.n, g.m, tiles0, kc_total, G, o.ncuts, max_pack, cx.arena.owned_bytes());
    std::random_device rd;
    std::vector<int32_t> best(G, 0);
    std::vector<uint8_t> have(G, 0);
    hipStream_t s = cx.stream;
    int32_t* gseg0 = nullptr;
    if (o.ncuts > 1) {
        gseg0 = cx.arena.get<int32_t>(g.n);
        std::vector<int32_t> gn(G), gv(G);
        for (int32_t i = 0; i < G; ++i) {
            gn[i] = gvtx[i + 1] - gvtx[i];
            gv[i] = gvtx[i];
        }
        fill_gseg(G, put(cx, gv), put(cx, gn), gseg0, s);
    }
    float initial_ms = 0.f;
    for (int32_t a0 = 0; a0 < o.ncuts; a0 += P) {
        const int32_t p = std::min<int32_t>(P, o.ncuts - a0);
        const Arena::Mark mark = cx.arena.mark();
        const int32_t VG = p * G;
        std::vector<int32_t> ugvtx((size_t)VG + 1), unparts(VG);
        std::vector<uint64_t> useeds(VG);
        for (int32_t j = 0; j < p; ++j)
            for (int32_t i = 0; i < G; ++i) {
                const int32_t vg = j * G + i;
                ugvtx[vg] = j * g.n + gvtx[i];
                unparts[vg] = nparts[i];
                uint64_t sd = seeds ? seeds[i] : o.seed;
                if (sd == 0) sd = ((uint64_t)rd() << 32) ^ rd();
                else sd += (uint64_t)(a0 + j);                                  // kway.rs:55-61
                useeds[vg] = sd;
            }
        ugvtx[VG] = p * g.n;
        DevGraph u = g;
        if (p > 1) {
            const int64_t un = (int64_t)p * g.n, um = (int64_t)p * g.m;
            u.n = (int32_t)un;
            u.m = (int32_t)um;
            u.xadj = cx.arena.get<int32_t>(un + 1);
            u.adjncy = cx.arena.get<int32_t>(um);
            u.adjwgt = g.adjwgt ? cx.arena.get<int32_t>(um) : nullptr;
            u.vwgt = g.vwgt ? cx.arena.get<int32_t>(un * g.ncon) : nullptr;
            u.vsize = nullptr;
            for (int32_t j = 0; j < p; ++j) {
                pack_csr(g.n, g.m, g.xadj, g.adjncy, j * g.n, j * g.m, u.xadj, u.adjncy, s);
                if (g.adjwgt)
                    RK_HIP(hipMemcpyAsync(u.adjwgt + (int64_t)j * g.m, g.adjwgt, sizeof(int32_t) * (size_t)g.m,
                                          hipMemcpyDeviceToDevice, s));
                if (g.vwgt)
                    RK_HIP(hipMemcpyAsync(u.vwgt + (int64_t)j * g.n * g.ncon, g.vwgt,
                                          sizeof(int32_t) * (size_t)g.n * g.ncon, hipMemcpyDeviceToDevice, s));
            }
            const int32_t* hend = put(cx, std::vector<int32_t>{(int32_t)um});
            RK_HIP(hipMemcpyAsync(u.xadj + un, hend, sizeof(int32_t), hipMemcpyDeviceToDevice, s));
        }
        int32_t* upart = (o.ncuts == 1) ? d_part : cx.arena.get<int32_t>(u.n);
        std::vector<int32_t> uobj(VG);
        batch_attempt(cx, u, VG, ugvtx.data(), unparts.data(), ubvec, o, useeds.data(), upart, uobj.data(), st);
        if (st) initial_ms += st->initial_ms;
        if (o.ncuts > 1) {
            std::vector<int32_t> choice(G, -1);
            for (int32_t j = 0; j < p; ++j)
                for (int32_t i = 0; i < G; ++i) {
                    const int32_t c = uobj[(size_t)j * G + i];
                    if (!have[i] || c < best[i]) {
                        have[i] = 1;
                        best[i] = c;
                        choice[i] = j;
                    }
                }
            copy_attempt(g.n, gseg0, put(cx, choice), upart, d_part, s);
            RK_HIP(hipStreamSynchronize(s));
        } else {
            for (int32_t i = 0; i < G; ++i) best[i] = uobj[i];
        }
        cx.arena.rewind(mark);
        cx.upload_reset();
    }
    for (int32_t i = 0; i < G; ++i) objval[i] = best[i];
    if (st) {
        st->attempts = o.ncuts;
        st->attempts_per_call = P;
        st->initial_ms = initial_ms;
    }
}

}  // namespace rk
