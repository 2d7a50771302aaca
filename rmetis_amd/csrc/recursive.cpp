// recursive.cpp -- recursive bisection (partition/recursive.rs:11-137) and
// nested dissection (partition/nd.rs:16-111 with separator/mod.rs:10-61) on
// the GPU.
//
// Both recurse over multilevel bisections of subgraphs.  Here the recursion
// runs breadth-first: at each depth every subgraph still to be split is one
// graph of a union in device memory, one batched call bisects all of them
// (bisect_batch_device: matching, contraction, growing and refinement of
// every subgraph in the same level kernels), and the union is split on the
// device into the next depth's subgraphs (split.hip), which are exactly the
// reference's extract_subgraph results.  Subgraphs that end the recursion
// are placed directly: all vertices to one part (recursive.rs:70-76), one
// vertex per part (79-86), or consecutive positions of the ordering
// (nd.rs:59-66, 77-83, 105-108).
//
// Recursive bisection carries the reference's target fractions exactly
// (recursive.rs:94-118, in f32).  Its random stream differs: the reference
// threads one generator through the depth-first recursion, while here each
// subgraph's bisection draws from its own seed, derived from its parent's
// and its side, so that a depth's bisections are independent.  Nested
// dissection re-seeds every bisection from the options' seed, as the
// reference's partition_recursive call does (recursive.rs:28-35), so the
// subgraphs of a depth share it.
#include <algorithm>
#include <climits>
#include <random>

#include "driver.hpp"

namespace rk {

namespace {

template <typename T>
T* up(Context& cx, const std::vector<T>& v, T* d = nullptr) {
    if (!d) d = cx.arena.get<T>((int64_t)v.size());
    if (!v.empty()) {
        T* h = (T*)cx.upload_stage(sizeof(T) * v.size());
        std::copy(v.begin(), v.end(), h);
        RK_HIP(hipMemcpyAsync(d, h, sizeof(T) * v.size(), hipMemcpyHostToDevice, cx.stream));
    }
    return d;
}

template <typename T>
std::vector<T> down(Context& cx, const T* d, size_t count) {
    std::vector<T> h(count);
    if (count) {
        T* st = (T*)cx.host_stage2(sizeof(T) * count);
        RK_HIP(hipMemcpyAsync(st, d, sizeof(T) * count, hipMemcpyDeviceToHost, cx.stream));
        RK_HIP(hipStreamSynchronize(cx.stream));
        std::copy(st, st + count, h.begin());
    }
    return h;
}

uint64_t entropy_seed() {
    std::random_device rd;
    return ((uint64_t)rd() << 32) ^ rd();
}

// A union of subgraphs in device memory.
struct Union {
    DevGraph g;
    const int32_t* orig = nullptr;   // vertex of the input graph (null: the input itself)
    std::vector<int32_t> gvtx;       // host: first vertex of each subgraph, then g.n
};

// Device buffers of the recursion, sized by the input graph (every depth's
// union is a subgraph of it): two unions used in turn, and the split tables.
struct Recursion {
    DevGraph buf[2];
    int32_t* orig[2] = {nullptr, nullptr};
    int32_t* part = nullptr;     // bisection of the current union
    int32_t* label = nullptr;
    int32_t* nid = nullptr;
    int32_t* rank = nullptr;
    int32_t* flag = nullptr;
    int32_t* pre[3] = {nullptr, nullptr, nullptr};
    const int32_t* const* d_pre = nullptr;
    int32_t* ndeg = nullptr;
    int32_t* gseg = nullptr;
    void* scan_tmp = nullptr;

    Recursion(Context& cx, const DevGraph& g, int L) {
        const int64_t n = g.n, m = g.m;
        for (auto& b : buf) {
            b.ncon = g.ncon;
            b.xadj = cx.arena.get<int32_t>(n + 1);
            b.adjncy = cx.arena.get<int32_t>(m);
            b.adjwgt = g.adjwgt ? cx.arena.get<int32_t>(m) : nullptr;
            b.vwgt = g.vwgt ? cx.arena.get<int32_t>(n * g.ncon) : nullptr;
        }
        orig[0] = cx.arena.get<int32_t>(n);
        orig[1] = cx.arena.get<int32_t>(n);
        part = cx.arena.get<int32_t>(n);
        label = cx.arena.get<int32_t>(n);
        nid = cx.arena.get<int32_t>(n);
        rank = cx.arena.get<int32_t>(n);
        flag = cx.arena.get<int32_t>(n);
        std::vector<int32_t*> p(3, nullptr);
        for (int l = 0; l < L; ++l) p[l] = pre[l] = cx.arena.get<int32_t>(n + 1);
        d_pre = (const int32_t* const*)up(cx, p);
        ndeg = cx.arena.get<int32_t>(n + 1);
        gseg = cx.arena.get<int32_t>(n);
        scan_tmp = cx.arena.alloc(scan_temp_bytes(n + 1));
    }
};

// The graph of each vertex of u (into R.gseg).
void union_segments(Context& cx, Recursion& R, const Union& u) {
    const int32_t G = (int32_t)u.gvtx.size() - 1;
    std::vector<int32_t> gn(G);
    for (int32_t i = 0; i < G; ++i) gn[i] = u.gvtx[i + 1] - u.gvtx[i];
    fill_gseg(G, up(cx, u.gvtx), up(cx, gn), R.gseg, cx.stream);
}

// Label prefixes of u and the per-subgraph counts of each label (host).
std::vector<int32_t> label_counts(Context& cx, Recursion& R, const Union& u, const int32_t* d_gvtx, int L) {
    const int32_t G = (int32_t)u.gvtx.size() - 1;
    for (int l = 0; l < L; ++l) label_prefix(u.g.n, R.label, l, R.flag, R.pre[l], R.scan_tmp, cx.stream);
    int32_t* d_counts = cx.arena.get<int32_t>((int64_t)G * L);
    split_counts(G, L, d_gvtx, R.d_pre, d_counts, cx.stream);
    return down(cx, d_counts, (size_t)G * L);
}

// Places the children with pa >= 0 and builds the union of the children with
// cbase >= 0 (nn vertices) into buffer `into`.
Union split_union(Context& cx, Recursion& R, const Union& u, const int32_t* d_gvtx, int L,
                  const std::vector<int32_t>& cbase, const std::vector<int32_t>& pa, const std::vector<int32_t>& pcap,
                  std::vector<int32_t> ngvtx, int32_t* out, int into) {
    hipStream_t s = cx.stream;
    split_rank(u.g.n, L, R.label, R.gseg, d_gvtx, R.d_pre, up(cx, cbase), R.nid, R.rank, s);
    place_children(u.g.n, L, R.label, R.gseg, R.rank, u.orig, up(cx, pa), up(cx, pcap), out, s);
    Union nu;
    nu.gvtx = std::move(ngvtx);
    const int32_t nn = nu.gvtx.empty() ? 0 : nu.gvtx.back();
    if (nn == 0) {
        nu.gvtx.clear();
        RK_HIP(hipStreamSynchronize(s));
        return nu;
    }
    nu.g = R.buf[into];
    nu.g.n = nn;
    split_graph(u.g, u.orig, R.label, R.nid, nn, nu.g, R.orig[into], R.ndeg, R.scan_tmp, s);
    nu.g.m = down(cx, nu.g.xadj + nn, 1)[0];
    nu.orig = R.orig[into];
    return nu;
}

}  // namespace

int32_t recursive_device(Context& cx, const DevGraph& g, int32_t nparts, const float* tpwgts, const float* ubvec,
                         const Options& o, int32_t* d_part) {
    if (nparts < 2) throw Error(-2, "nparts must be >= 2, got " + std::to_string(nparts));   // recursive.rs:19-21
    if (nparts > g.n) throw Error(-2, "nparts > nvtxs");                                    // recursive.rs:22-26
    hipStream_t s = cx.stream;
    const int32_t ncon = g.ncon;
    const uint64_t seed0 = o.seed == 0 ? entropy_seed() : o.seed;                            // recursive.rs:28-35
    struct Sub {
        int32_t k, off;
        std::vector<float> tp;   // the k parts' target fractions [k * ncon]
        uint64_t seed;
    };
    std::vector<Sub> subs(1);
    subs[0].k = nparts;
    subs[0].off = 0;
    subs[0].tp = tpwgts ? std::vector<float>(tpwgts, tpwgts + (size_t)nparts * ncon)        // recursive.rs:37-43
                        : std::vector<float>((size_t)nparts * ncon, 1.0f / (float)nparts);
    subs[0].seed = seed0;
    Recursion R(cx, g, 2);
    Union u;
    u.g = g;
    u.gvtx = {0, g.n};
    if (g.n <= nparts) {                                                 // recursive.rs:79-86 (n == nparts)
        RK_HIP(hipMemsetAsync(R.label, 0, sizeof(int32_t) * (size_t)g.n, s));
        union_segments(cx, R, u);
        const int32_t* d_gvtx = up(cx, u.gvtx);
        label_counts(cx, R, u, d_gvtx, 2);
        split_union(cx, R, u, d_gvtx, 2, {-1, -1}, {0, -1}, {nparts - 1, 0}, {}, d_part, 0);
        subs.clear();
    }
    int into = 0;
    while (!subs.empty()) {
        const int32_t G = (int32_t)subs.size();
        const Arena::Mark mark = cx.arena.mark();
        union_segments(cx, R, u);
        const int32_t* d_gvtx = up(cx, u.gvtx);
        // the bisection targets of every subgraph (recursive.rs:89-105)
        std::vector<float> tp2((size_t)2 * ncon * G);
        std::vector<std::vector<float>> lw(G, std::vector<float>(ncon)), rw(G, std::vector<float>(ncon));
        std::vector<uint64_t> seeds(G);
        for (int32_t i = 0; i < G; ++i) {
            const Sub& sb = subs[i];
            const int32_t k1 = sb.k / 2;
            for (int32_t c = 0; c < ncon; ++c) {
                float l = 0.f, r = 0.f;
                for (int32_t p = 0; p < k1; ++p) l += sb.tp[(size_t)p * ncon + c];
                for (int32_t p = k1; p < sb.k; ++p) r += sb.tp[(size_t)p * ncon + c];
                lw[i][c] = l;
                rw[i][c] = r;
            }
            const float total = lw[i][0] + rw[i][0];
            for (int32_t c = 0; c < ncon; ++c) {
                tp2[(size_t)2 * ncon * i + c] = lw[i][c] / total;
                tp2[(size_t)2 * ncon * i + ncon + c] = rw[i][c] / total;
            }
            seeds[i] = sb.seed;
        }
        bisect_batch_device(cx, u.g, G, u.gvtx.data(), tp2.data(), ubvec, o, seeds.data(), R.part);   // recursive.rs:107
        RK_HIP(hipMemcpyAsync(R.label, R.part, sizeof(int32_t) * (size_t)u.g.n, hipMemcpyDeviceToDevice, s));
        const std::vector<int32_t> cnt = label_counts(cx, R, u, d_gvtx, 2);
        // the children (recursive.rs:110-132): (i, 0) takes parts off..off+k1,
        // (i, 1) the rest, with the fractions renormalised by the side's sum
        std::vector<int32_t> cbase(2 * (size_t)G, -1), pa(2 * (size_t)G, -1), pcap(2 * (size_t)G, 0);
        std::vector<int32_t> ngvtx{0};
        std::vector<Sub> next;
        for (int32_t i = 0; i < G; ++i) {
            const Sub& sb = subs[i];
            const int32_t k1 = sb.k / 2;
            for (int l = 0; l < 2; ++l) {
                const int32_t cn = cnt[(size_t)2 * i + l];
                if (cn == 0) continue;
                Sub c;
                c.k = l == 0 ? k1 : sb.k - k1;
                c.off = sb.off + (l == 0 ? 0 : k1);
                const float sum = l == 0 ? lw[i][0] : rw[i][0];
                c.tp.assign(sb.tp.begin() + (l == 0 ? 0 : (size_t)k1 * ncon),
                            l == 0 ? sb.tp.begin() + (size_t)k1 * ncon : sb.tp.end());
                for (float& x : c.tp) x /= sum;
                uint64_t st = sb.seed + (uint64_t)l + 1;
                c.seed = Rng::splitmix(st);
                const size_t ci = (size_t)2 * i + l;
                if (c.k == 1 || cn <= c.k) {                             // recursive.rs:70-86
                    pa[ci] = c.off;
                    pcap[ci] = c.k - 1;
                } else {
                    cbase[ci] = ngvtx.back();
                    ngvtx.push_back(ngvtx.back() + cn);
                    next.push_back(std::move(c));
                }
            }
        }
        if (next.empty()) ngvtx.clear();
        Union nu = split_union(cx, R, u, d_gvtx, 2, cbase, pa, pcap, ngvtx, d_part, into);
        RK_HIP(hipStreamSynchronize(s));
        cx.arena.rewind(mark);
        cx.upload_reset();
        u = std::move(nu);
        subs = std::move(next);
        into ^= 1;
    }
    // objval = edge cut of the whole partition (recursive.rs:55)
    unsigned long long* acc = cx.arena.get<unsigned long long>(1);
    int32_t* cut = cx.arena.get<int32_t>(1);
    edge_cut(g, d_part, acc, cut, pick_group(g.n, g.m), s);
    return down(cx, cut, 1)[0];
}

void node_nd_device(Context& cx, const DevGraph& g, const Options& o, int32_t* d_perm, int32_t* d_iperm) {
    constexpr int32_t kNdThreshold = 100;                                // nd.rs:13
    hipStream_t s = cx.stream;
    Recursion R(cx, g, 3);
    Union u;
    u.g = g;
    u.gvtx = {0, g.n};
    std::vector<int32_t> start{0};                                       // first position of each subgraph
    if (g.n <= kNdThreshold) {                                           // nd.rs:59-66
        RK_HIP(hipMemsetAsync(R.label, 0, sizeof(int32_t) * (size_t)g.n, s));
        union_segments(cx, R, u);
        const int32_t* d_gvtx = up(cx, u.gvtx);
        label_counts(cx, R, u, d_gvtx, 3);
        split_union(cx, R, u, d_gvtx, 3, {-1, -1, -1}, {0, -1, -1}, {INT32_MAX, 0, 0}, {}, d_perm, 0);
        start.clear();
    }
    // every bisection is partition_recursive(graph, 2, [0.5, 0.5], [1.05])
    // (separator/mod.rs:19-28), re-seeded from the options (recursive.rs:28-35)
    const float ub[1] = {1.05f};
    int into = 0;
    while (!start.empty()) {
        const int32_t G = (int32_t)start.size();
        const Arena::Mark mark = cx.arena.mark();
        union_segments(cx, R, u);
        const int32_t* d_gvtx = up(cx, u.gvtx);
        std::vector<float> tp2(2 * (size_t)G, 0.5f);
        std::vector<uint64_t> seeds(G);
        for (int32_t i = 0; i < G; ++i) seeds[i] = o.seed == 0 ? entropy_seed() : o.seed;
        bisect_batch_device(cx, u.g, G, u.gvtx.data(), tp2.data(), ub, o, seeds.data(), R.part);
        // the separator: the lighter side's vertices with a neighbour across
        // (separator/mod.rs:29-58)
        std::vector<int32_t> gk(G, 2), gkc(G);
        for (int32_t i = 0; i < G; ++i) gkc[i] = 2 * i;
        int32_t* d_pw = cx.arena.get<int32_t>(2 * (int64_t)G);
        part_weights_batch(u.g, R.part, R.gseg, up(cx, gkc), up(cx, gk), 2 * (int64_t)G, d_pw, s);
        separator_labels(u.g, R.part, R.gseg, d_pw, R.label, s);
        std::vector<int32_t> cnt = label_counts(cx, R, u, d_gvtx, 3);
        // a separator of half the vertices or more orders the subgraph as it
        // is (nd.rs:77-83): all of its vertices become label 0, placed in order
        std::vector<uint8_t> degen(G, 0);
        bool anydegen = false;
        for (int32_t i = 0; i < G; ++i) {
            const int32_t n = u.gvtx[i + 1] - u.gvtx[i];
            if (cnt[(size_t)3 * i + 2] >= n / 2) {
                degen[i] = 1;
                anydegen = true;
            }
        }
        if (anydegen) {
            relabel_graphs(u.g.n, R.gseg, up(cx, degen), 0, R.label, s);
            label_prefix(u.g.n, R.label, 0, R.flag, R.pre[0], R.scan_tmp, s);
            label_prefix(u.g.n, R.label, 2, R.flag, R.pre[2], R.scan_tmp, s);
            label_prefix(u.g.n, R.label, 1, R.flag, R.pre[1], R.scan_tmp, s);
        }
        std::vector<int32_t> cbase(3 * (size_t)G, -1), pa(3 * (size_t)G, -1), pcap(3 * (size_t)G, INT32_MAX);
        std::vector<int32_t> ngvtx{0}, nstart;
        for (int32_t i = 0; i < G; ++i) {
            const int32_t n = u.gvtx[i + 1] - u.gvtx[i];
            if (degen[i]) {
                pa[(size_t)3 * i] = start[i];
                continue;
            }
            const int32_t nl = cnt[(size_t)3 * i], nr = cnt[(size_t)3 * i + 1];
            pa[(size_t)3 * i + 2] = start[i] + nl + nr;                  // nd.rs:105-108
            const int32_t cs[2] = {start[i], start[i] + nl};            // left, then right (nd.rs:92-102)
            const int32_t cn[2] = {nl, nr};
            for (int l = 0; l < 2; ++l) {
                if (cn[l] == 0) continue;
                const size_t ci = (size_t)3 * i + l;
                if (cn[l] <= kNdThreshold) {
                    pa[ci] = cs[l];
                } else {
                    cbase[ci] = ngvtx.back();
                    ngvtx.push_back(ngvtx.back() + cn[l]);
                    nstart.push_back(cs[l]);
                }
            }
            (void)n;
        }
        if (nstart.empty()) ngvtx.clear();
        Union nu = split_union(cx, R, u, d_gvtx, 3, cbase, pa, pcap, ngvtx, d_perm, into);
        RK_HIP(hipStreamSynchronize(s));
        cx.arena.rewind(mark);
        cx.upload_reset();
        u = std::move(nu);
        start = std::move(nstart);
        into ^= 1;
    }
    inverse_permutation(g.n, d_perm, d_iperm, s);                        // nd.rs:38-41
    RK_HIP(hipStreamSynchronize(s));
}

}  // namespace rk
