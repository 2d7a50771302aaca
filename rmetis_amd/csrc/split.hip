// split.hip -- subgraph extraction and placement for recursive bisection and
// nested dissection on a union of graphs.
//
// Graph::extract_subgraph (graph/mod.rs:209-270) keeps the vertices of one
// part in ascending order and the edges between them in CSR order, with
// their weights.  Here every graph of a union is split at once: a vertex of
// graph g with label l >= 0 goes to child (g, l); labels below 0 leave the
// union.  A child's vertices keep their order, and the children are laid out
// one after another, so each child is exactly the reference's subgraph.
//
//  * k_split_rank: per vertex its child-local rank (from per-label prefix
//    counts) and, for children that stay, its new id;
//  * k_split_degree / k_split_fill: the kept degree of each new vertex, then
//    (after a scan) its kept neighbours (renumbered) and weights, the vertex
//    weights and the original-vertex map;
//  * k_place: writes out[orig[v]] = a + min(rank, cap) for the vertices of
//    children that end here (a recursion leaf: recursive.rs:70-86, nd.rs:59-66).
// A null orig map is the identity (the input graph itself).
#include <climits>

#include "common.hpp"
#include "kernels.hpp"

namespace rk {

namespace {

constexpr int kSplitG = 8;   // lanes per vertex in the degree and fill kernels

__global__ __launch_bounds__(kBlock) void k_label_flags(int32_t n, const int32_t* __restrict__ label, int32_t l,
                                                        int32_t* __restrict__ flag) {
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock)
        flag[v] = label[v] == l ? 1 : 0;
}

// Per graph g and label l < L: the count of its label-l vertices, from the
// exclusive prefix counts pre[l] (n + 1 entries each).
__global__ __launch_bounds__(kBlock) void k_split_counts(int32_t ngraphs, int32_t L, const int32_t* __restrict__ gv0,
                                                         const int32_t* const* __restrict__ pre,
                                                         int32_t* __restrict__ counts) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < (int64_t)ngraphs * L;
         i += (int64_t)gridDim.x * kBlock) {
        const int32_t g = (int32_t)(i / L), l = (int32_t)(i % L);
        counts[i] = pre[l][gv0[g + 1]] - pre[l][gv0[g]];
    }
}

// nid[v] = the new id of v in the child union (-1 when its child leaves the
// union); rank[v] = v's position within its child (-1 without a label).
__global__ __launch_bounds__(kBlock) void k_split_rank(int32_t n, int32_t L, const int32_t* __restrict__ label,
                                                       const int32_t* __restrict__ gseg,
                                                       const int32_t* __restrict__ gv0,
                                                       const int32_t* const* __restrict__ pre,
                                                       const int32_t* __restrict__ cbase, int32_t* __restrict__ nid,
                                                       int32_t* __restrict__ rank) {
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock) {
        const int32_t l = label[v];
        int32_t r = -1, id = -1;
        if (l >= 0) {
            const int32_t g = gseg[v];
            r = pre[l][v] - pre[l][gv0[g]];
            const int32_t b = cbase[(int64_t)g * L + l];
            if (b >= 0) id = b + r;
        }
        nid[v] = id;
        rank[v] = r;
    }
}

// ndeg[nid[v]] = the number of v's neighbours with v's label.
__global__ __launch_bounds__(kBlock) void k_split_degree(int32_t n, const int32_t* __restrict__ xadj,
                                                         const int32_t* __restrict__ adjncy,
                                                         const int32_t* __restrict__ label,
                                                         const int32_t* __restrict__ nid, int32_t* __restrict__ ndeg) {
    const int gl = group_lane<kSplitG>();
    for (int64_t v = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kSplitG; v < n;
         v += (int64_t)gridDim.x * (kBlock / kSplitG)) {
        const int32_t id = nid[v];
        if (id < 0) continue;                                   // group-uniform
        const int32_t l = label[v];
        int32_t c = 0;
        for (int32_t j = xadj[v] + gl; j < xadj[v + 1]; j += kSplitG) c += label[adjncy[j]] == l ? 1 : 0;
        c = group_sum<kSplitG>(c);
        if (gl == 0) ndeg[id] = c;
    }
}

// The kept neighbours of v in CSR order (graph/mod.rs:237-250), renumbered,
// with their weights; v's vertex weights and original id.
__global__ __launch_bounds__(kBlock) void k_split_fill(int32_t n, int32_t ncon, const int32_t* __restrict__ xadj,
                                                       const int32_t* __restrict__ adjncy,
                                                       const int32_t* __restrict__ adjwgt,
                                                       const int32_t* __restrict__ vwgt,
                                                       const int32_t* __restrict__ orig,
                                                       const int32_t* __restrict__ label,
                                                       const int32_t* __restrict__ nid,
                                                       const int32_t* __restrict__ nxadj, int32_t* __restrict__ nadj,
                                                       int32_t* __restrict__ nw, int32_t* __restrict__ nvw,
                                                       int32_t* __restrict__ norig) {
    const int gl = group_lane<kSplitG>();
    for (int64_t v = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kSplitG; v < n;
         v += (int64_t)gridDim.x * (kBlock / kSplitG)) {
        const int32_t id = nid[v];
        if (id < 0) continue;                                   // group-uniform
        const int32_t l = label[v];
        int32_t out = nxadj[id];
        const int32_t beg = xadj[v], end = xadj[v + 1];
        for (int32_t j0 = beg; j0 < end; j0 += kSplitG) {
            const int32_t j = j0 + gl;
            int32_t u = 0;
            bool keep = false;
            if (j < end) {
                u = adjncy[j];
                keep = label[u] == l;
            }
            const uint64_t m = group_ballot<kSplitG>(keep);
            if (keep) {
                const int32_t at = out + __popcll(m & ((1ull << gl) - 1ull));
                nadj[at] = nid[u];
                if (nw) nw[at] = adjwgt[j];
            }
            out += __popcll(m);
        }
        if (gl == 0) {
            norig[id] = orig ? orig[v] : (int32_t)v;
            if (nvw)
                for (int c = 0; c < ncon; ++c) nvw[(int64_t)id * ncon + c] = vwgt[v * (int64_t)ncon + c];
        }
    }
}

// out[orig[v]] = pa[c] + min(rank[v], pcap[c]) for v of child c = g*L +
// label[v] with pa[c] >= 0.
__global__ __launch_bounds__(kBlock) void k_place(int32_t n, int32_t L, const int32_t* __restrict__ label,
                                                  const int32_t* __restrict__ gseg, const int32_t* __restrict__ rank,
                                                  const int32_t* __restrict__ orig, const int32_t* __restrict__ pa,
                                                  const int32_t* __restrict__ pcap, int32_t* __restrict__ out) {
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock) {
        const int32_t l = label[v];
        if (l < 0) continue;
        const int64_t c = (int64_t)gseg[v] * L + l;
        const int32_t a = pa[c];
        if (a < 0) continue;
        out[orig ? orig[v] : v] = a + min(rank[v], pcap[c]);
    }
}

// label[v] = l for the vertices of the graphs g with flag[g] != 0.
__global__ __launch_bounds__(kBlock) void k_relabel(int32_t n, const int32_t* __restrict__ gseg,
                                                    const uint8_t* __restrict__ flag, int32_t l,
                                                    int32_t* __restrict__ label) {
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock)
        if (flag[gseg[v]]) label[v] = l;
}

__global__ __launch_bounds__(kBlock) void k_inverse(int32_t n, const int32_t* __restrict__ perm,
                                                    int32_t* __restrict__ iperm) {
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock)
        iperm[perm[v]] = (int32_t)v;
}

// Vertex separator of a bisected graph (separator/mod.rs:29-58): side[v] =
// the bisection part; smaller[g] = the lighter side on constraint 0 (pw0 <=
// pw1 gives side 0).  label[v] = 2 when v is on the smaller side with a
// neighbour on the other, else 0 for the smaller side and 1 for the other.
__global__ __launch_bounds__(kBlock) void k_separator(int32_t n, const int32_t* __restrict__ xadj,
                                                      const int32_t* __restrict__ adjncy,
                                                      const int32_t* __restrict__ side,
                                                      const int32_t* __restrict__ gseg, const int32_t* __restrict__ pw,
                                                      int32_t ncon, int32_t* __restrict__ label) {
    const int gl = group_lane<kSplitG>();
    for (int64_t v = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kSplitG; v < n;
         v += (int64_t)gridDim.x * (kBlock / kSplitG)) {
        const int32_t g = gseg[v];
        const int32_t* p = pw + (int64_t)g * 2 * ncon;
        const int32_t smaller = p[0] <= p[ncon] ? 0 : 1;
        const int32_t sv = side[v];
        bool cut = false;
        if (sv == smaller)
            for (int32_t j = xadj[v] + gl; j < xadj[v + 1] && !cut; j += kSplitG) cut = side[adjncy[j]] != sv;
        cut = group_ballot<kSplitG>(cut) != 0;
        if (gl == 0) label[v] = cut ? 2 : (sv == smaller ? 0 : 1);
    }
}

inline int grid_for(int64_t items, int per_block) {
    int64_t b = (items + per_block - 1) / per_block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 16384));
}

}  // namespace

void label_prefix(int32_t n, const int32_t* label, int32_t l, int32_t* flag, int32_t* pre, void* scan_tmp,
                  hipStream_t s) {
    hipLaunchKernelGGL(k_label_flags, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, label, l, flag);
    exclusive_scan_i32(flag, pre, n, pre + n, scan_tmp, s);
    RK_HIP(hipGetLastError());
}

void split_counts(int32_t ngraphs, int32_t L, const int32_t* gv0, const int32_t* const* pre, int32_t* counts,
                  hipStream_t s) {
    hipLaunchKernelGGL(k_split_counts, dim3(grid_for((int64_t)ngraphs * L, kBlock)), dim3(kBlock), 0, s, ngraphs, L,
                       gv0, pre, counts);
    RK_HIP(hipGetLastError());
}

void split_rank(int32_t n, int32_t L, const int32_t* label, const int32_t* gseg, const int32_t* gv0,
                const int32_t* const* pre, const int32_t* cbase, int32_t* nid, int32_t* rank, hipStream_t s) {
    hipLaunchKernelGGL(k_split_rank, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, L, label, gseg, gv0, pre,
                       cbase, nid, rank);
    RK_HIP(hipGetLastError());
}

void split_graph(const DevGraph& g, const int32_t* orig, const int32_t* label, const int32_t* nid, int32_t nn,
                 DevGraph& out, int32_t* norig, int32_t* ndeg, void* scan_tmp, hipStream_t s) {
    const int grid = grid_for(g.n, kBlock / kSplitG);
    hipLaunchKernelGGL(k_split_degree, dim3(grid), dim3(kBlock), 0, s, g.n, g.xadj, g.adjncy, label, nid, ndeg);
    exclusive_scan_i32(ndeg, out.xadj, nn, out.xadj + nn, scan_tmp, s);
    hipLaunchKernelGGL(k_split_fill, dim3(grid), dim3(kBlock), 0, s, g.n, g.ncon, g.xadj, g.adjncy, g.adjwgt, g.vwgt,
                       orig, label, nid, out.xadj, out.adjncy, out.adjwgt, out.vwgt, norig);
    RK_HIP(hipGetLastError());
}

void place_children(int32_t n, int32_t L, const int32_t* label, const int32_t* gseg, const int32_t* rank,
                    const int32_t* orig, const int32_t* pa, const int32_t* pcap, int32_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_place, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, L, label, gseg, rank, orig, pa,
                       pcap, out);
    RK_HIP(hipGetLastError());
}

void relabel_graphs(int32_t n, const int32_t* gseg, const uint8_t* flag, int32_t l, int32_t* label, hipStream_t s) {
    hipLaunchKernelGGL(k_relabel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, gseg, flag, l, label);
    RK_HIP(hipGetLastError());
}

void inverse_permutation(int32_t n, const int32_t* perm, int32_t* iperm, hipStream_t s) {
    hipLaunchKernelGGL(k_inverse, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, perm, iperm);
    RK_HIP(hipGetLastError());
}

void separator_labels(const DevGraph& g, const int32_t* side, const int32_t* gseg, const int32_t* pw,
                      int32_t* label, hipStream_t s) {
    hipLaunchKernelGGL(k_separator, dim3(grid_for(g.n, kBlock / kSplitG)), dim3(kBlock), 0, s, g.n, g.xadj,
                       g.adjncy, side, gseg, pw, g.ncon, label);
    RK_HIP(hipGetLastError());
}

}  // namespace rk
