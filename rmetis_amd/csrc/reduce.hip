// reduce.hip -- edge-cut and part-weight reductions, partition projection.
//
// edge cut:      graph/mod.rs:123-137 (sum over directed entries, i32 wrap, /2)
// part weights:  graph/mod.rs:140-149 (pw[part[v]][c] += vwgt[v][c], i32 wrap)
// projection:    partition/kway.rs:81-84 (fine_part[v] = part[cmap[v]])
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace rk {

namespace {

constexpr int kCutWaves = 4;              // waves (independent tiles) per edge-cut block
constexpr int kCutCap = 1024;             // entries a wave stages per batch
constexpr int kCutItems = kCutCap / 64;
constexpr int kCutBlocksPerCU = 4;

// Edge cut, staged like the gain kernel: every wave takes tiles of 64
// consecutive vertices (lane = vertex, XCD-contiguous tile regions).  A
// batch of whole vertices is one contiguous entry range: neighbour ids and
// weights by range-checked buffer loads, then all part[] gathers in flight,
// then (part, weight) pairs in LDS; each lane sums the weights of its entries
// into other parts.  A tile holding a vertex of more than kCutCap entries is
// summed directly (a lane per vertex, the whole wave for each such vertex).
// Sums are exact in 64 bits and added once per wave (order independent).
__global__ __launch_bounds__(64 * kCutWaves) void k_edge_cut_tile(int32_t n, const int32_t* __restrict__ xadj,
                                                                  const int32_t* __restrict__ adjncy,
                                                                  const int32_t* __restrict__ adjwgt,
                                                                  const int32_t* __restrict__ part,
                                                                  unsigned long long* __restrict__ acc) {
    __shared__ int2 s_e_all[kCutWaves][kCutCap];
    const int wv = threadIdx.x >> 6, lane = __lane_id();
    int2* s_e = s_e_all[wv];
    const int ntiles = (int)(((int64_t)n + 63) / 64);
    const WaveTiles wt = wave_tiles(ntiles, kCutWaves);
    unsigned long long sum = 0;
    auto w_at = [&](int64_t j) __attribute__((always_inline)) -> uint32_t {
        return adjwgt ? (uint32_t)adjwgt[j] : 1u;
    };
    for (int it = 0; it < wt.count; ++it) {
        const int tile = wt.base + it * wt.stride;
        const int64_t v = (int64_t)tile * 64 + lane;
        const bool own = v < n;
        int32_t beg = 0, end = 0, pv = -1;
        if (own) {
            beg = xadj[v];
            end = xadj[v + 1];
            pv = part[v];
        }
        const int32_t deg = end - beg;
        uint64_t bigm = __ballot(deg > kCutCap);
        if (bigm) {
            if (deg <= kCutCap)
                for (int32_t j = beg; j < end; ++j)
                    if (part[adjncy[j]] != pv) sum += w_at(j);
            while (bigm) {
                const int b = __builtin_ctzll(bigm);
                bigm &= bigm - 1;
                const int32_t bb = __shfl(beg, b, 64), be = __shfl(end, b, 64), bp = __shfl(pv, b, 64);
                for (int32_t j = bb + lane; j < be; j += 64)
                    if (part[adjncy[j]] != bp) sum += w_at(j);
            }
            continue;
        }
        // exclusive prefix of the degrees over the tile
        int inc = deg;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(inc, o, 64);
            if (lane >= o) inc += y;
        }
        const int loff = inc - deg;
        const int32_t x0 = __builtin_amdgcn_readfirstlane(beg);   // lane 0 owns v0 < n
        for (int vs = 0; vs < 64;) {
            const int P0 = __shfl(loff, vs, 64);
            const uint64_t fit = __ballot(lane >= vs && loff + deg - P0 <= kCutCap);
            const int ve = 64 - __builtin_clzll(fit);               // fit is lanes [vs, ve), ve > vs
            const int P1 = __shfl(loff + deg, ve - 1, 64);
            if (P1 > P0) {
                const int p0 = __builtin_amdgcn_readfirstlane(P0);
                const int ne = __builtin_amdgcn_readfirstlane(P1 - P0);
                const int64_t e0 = (int64_t)x0 + p0;
                const int32_t* wsrc = adjwgt ? adjwgt : adjncy;
                const int lb4 = lane * 4;
                int32_t uu[kCutItems], ww[kCutItems], pu[kCutItems];
#pragma unroll
                for (int i = 0; i < kCutItems; ++i)
                    uu[i] = __builtin_amdgcn_raw_buffer_load_b32(edge_rsrc(adjncy + e0 + i * 64, max(ne - i * 64, 0)),
                                                                 lb4, 0, 0);
#pragma unroll
                for (int i = 0; i < kCutItems; ++i)
                    ww[i] = __builtin_amdgcn_raw_buffer_load_b32(edge_rsrc(wsrc + e0 + i * 64, max(ne - i * 64, 0)),
                                                                 lb4, 0, 0);
#pragma unroll
                for (int i = 0; i < kCutItems; ++i) pu[i] = part[uu[i]];
#pragma unroll
                for (int i = 0; i < kCutItems; ++i) {
                    const int p = P0 + lane + i * 64;
                    if (p < P1) s_e[p - P0] = make_int2(pu[i], adjwgt ? ww[i] : 1);
                }
            }
            wave_sync();
            if (lane >= vs && lane < ve && deg > 0) {
                const int lb = loff - P0, le = lb + deg;
                int e = lb;
                for (; e + 4 <= le; e += 4) {
                    const int2 a0 = s_e[e], a1 = s_e[e + 1], a2 = s_e[e + 2], a3 = s_e[e + 3];
                    sum += (a0.x != pv ? (uint32_t)a0.y : 0u) + (uint64_t)(a1.x != pv ? (uint32_t)a1.y : 0u) +
                           (uint64_t)(a2.x != pv ? (uint32_t)a2.y : 0u) + (uint64_t)(a3.x != pv ? (uint32_t)a3.y : 0u);
                }
                for (; e < le; ++e) {
                    const int2 a0 = s_e[e];
                    if (a0.x != pv) sum += (uint32_t)a0.y;
                }
            }
            wave_sync();
            vs = ve;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
    if (lane == 0 && sum) atomicAdd(acc, sum);
}

constexpr int kPwVerts = 4096;   // vertices per block in the part-weight kernel
constexpr int kPwLds = 8192;     // max k*ncon histogram entries kept in LDS

__global__ __launch_bounds__(kBlock) void k_part_weights_lds(int32_t n, int32_t ncon,
                                                             const int32_t* __restrict__ vwgt,
                                                             const int32_t* __restrict__ part,
                                                             int32_t kc, int32_t* __restrict__ pw) {
    __shared__ int32_t s_h[kPwLds];
    for (int i = threadIdx.x; i < kc; i += kBlock) s_h[i] = 0;
    __syncthreads();
    const int64_t beg = (int64_t)blockIdx.x * kPwVerts;
    const int64_t end = min((int64_t)n, beg + kPwVerts);
    for (int64_t v = beg + threadIdx.x; v < end; v += kBlock) {
        const int32_t p = part[v];
        for (int c = 0; c < ncon; ++c)
            atomicAdd(&s_h[p * ncon + c], vwgt ? vwgt[v * ncon + c] : 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kc; i += kBlock)
        if (s_h[i] != 0) atomicAdd(&pw[i], s_h[i]);
}

__global__ __launch_bounds__(kBlock) void k_part_weights_global(int32_t n, int32_t ncon,
                                                                const int32_t* __restrict__ vwgt,
                                                                const int32_t* __restrict__ part,
                                                                int32_t* __restrict__ pw) {
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock) {
        const int32_t p = part[v];
        for (int c = 0; c < ncon; ++c) atomicAdd(&pw[(int64_t)p * ncon + c], vwgt ? vwgt[v * ncon + c] : 1);
    }
}

// Vertices with cmap -1 (graphs of a batch whose coarsest level this is)
// keep the part already stored for them.  Block 0 also clears clear_a[0..na)
// and clear_b[0..nb) for the level's refinement (part weights, done flag)
// in place of separate fills.
__global__ __launch_bounds__(kBlock) void k_project(int32_t n, const int32_t* __restrict__ cmap,
                                                    const int32_t* __restrict__ part_c,
                                                    int32_t* __restrict__ part_f, int32_t* __restrict__ clear_a,
                                                    int64_t na, int32_t* __restrict__ clear_b, int32_t nb) {
    int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v < n) {
        const int32_t c = cmap[v];
        if (c >= 0) part_f[v] = part_c[c];
    }
    if (blockIdx.x == 0) {
        for (int64_t i = threadIdx.x; i < na; i += kBlock) clear_a[i] = 0;
        for (int32_t i = threadIdx.x; i < nb; i += kBlock) clear_b[i] = 0;
    }
}

// ---- batched levels: several graphs laid out one after another ----

// pw of every graph at its table offset gkc[g], or (gmask non-null) of the
// graphs with gmask[g] set only, whose tables start at zero; a block whose
// vertices all belong to one graph sums in LDS first.
__global__ __launch_bounds__(kBlock) void k_part_weights_batch(int32_t n, int32_t ncon,
                                                               const int32_t* __restrict__ vwgt,
                                                               const int32_t* __restrict__ part,
                                                               const int32_t* __restrict__ gseg,
                                                               const int32_t* __restrict__ gkc,
                                                               const int32_t* __restrict__ gk,
                                                               const uint8_t* __restrict__ gmask,
                                                               int32_t* __restrict__ pw) {
    __shared__ int32_t s_h[kPwLds];
    const int64_t beg = (int64_t)blockIdx.x * kPwVerts;
    const int64_t end = min((int64_t)n, beg + kPwVerts);
    const int32_t g0 = gseg[beg], g1 = gseg[end - 1];
    const int kc = gk[g0] * ncon;
    if (g0 == g1 && kc <= kPwLds) {
        if (gmask && !gmask[g0]) return;
        for (int i = threadIdx.x; i < kc; i += kBlock) s_h[i] = 0;
        __syncthreads();
        for (int64_t v = beg + threadIdx.x; v < end; v += kBlock) {
            const int32_t p = part[v];
            for (int c = 0; c < ncon; ++c) atomicAdd(&s_h[p * ncon + c], vwgt ? vwgt[v * ncon + c] : 1);
        }
        __syncthreads();
        int32_t* out = pw + gkc[g0];
        for (int i = threadIdx.x; i < kc; i += kBlock)
            if (s_h[i] != 0) atomicAdd(&out[i], s_h[i]);
    } else {
        for (int64_t v = beg + threadIdx.x; v < end; v += kBlock) {
            const int32_t g = gseg[v];
            if (gmask && !gmask[g]) continue;
            int32_t* out = pw + gkc[g];
            const int32_t p = part[v];
            for (int c = 0; c < ncon; ++c) atomicAdd(&out[p * ncon + c], vwgt ? vwgt[v * ncon + c] : 1);
        }
    }
}

// graph/validate.rs:16-142 on the device for a graph already uploaded whose
// xadj starts at 0 and never decreases (checked on the host): every
// neighbour is in range and not the vertex itself (rows are read only
// inside [0, m)), and every vertex appears in as many rows as its own row is
// long; then vertex weights are positive.  *bad is set by any violation.
__global__ __launch_bounds__(kBlock) void k_validate_rows(int32_t n, int64_t m, const int32_t* __restrict__ xadj,
                                                          const int32_t* __restrict__ adjncy,
                                                          int32_t* __restrict__ cnt, int32_t* __restrict__ bad) {
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock) {
        const int64_t b = xadj[v], e = xadj[v + 1];
        if (b < 0 || b > e || e > m) {
            *bad = 1;
            continue;
        }
        for (int64_t j = b; j < e; ++j) {
            const int32_t u = adjncy[j];
            if (u < 0 || u >= n || u == (int32_t)v) {
                *bad = 1;
                break;
            }
            atomicAdd(&cnt[u], 1);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_validate_counts(int32_t n, int32_t ncon, const int32_t* __restrict__ xadj,
                                                            const int32_t* __restrict__ vwgt,
                                                            const int32_t* __restrict__ cnt, int32_t* __restrict__ bad) {
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock) {
        if (cnt[v] != xadj[v + 1] - xadj[v]) *bad = 1;
        if (vwgt)
            for (int c = 0; c < ncon; ++c)
                if (vwgt[v * ncon + c] <= 0) *bad = 1;
    }
}

// Zeroes the part-weight tables of the graphs with gmask[g] set (block g).
__global__ __launch_bounds__(kBlock) void k_zero_masked(int32_t ngraphs, int32_t ncon, const int32_t* __restrict__ gkc,
                                                        const int32_t* __restrict__ gk,
                                                        const uint8_t* __restrict__ gmask, int32_t* __restrict__ pw) {
    for (int32_t g = blockIdx.x; g < ngraphs; g += gridDim.x) {
        if (!gmask[g]) continue;
        int32_t* out = pw + gkc[g];
        const int kc = gk[g] * ncon;
        for (int i = threadIdx.x; i < kc; i += kBlock) out[i] = 0;
    }
}

// Edge cut of every graph.  A wave takes 64 consecutive vertices, whose rows
// form one contiguous span of entries: lane q of a round reads span position
// q (coalesced adjncy/adjwgt loads, then the part[] gather), its vertex found
// by a binary search over the 64 row starts in LDS.  Sums wrap in 32 bits
// (only the low 32 bits of a graph's total are used).  When the 64 vertices
// belong to one graph the wave adds its sum with one atomic; otherwise each
// vertex's sum is gathered in LDS and added to its own graph.
__global__ __launch_bounds__(kBlock) void k_edge_cut_batch(int32_t n, const int32_t* __restrict__ xadj,
                                                           const int32_t* __restrict__ adjncy,
                                                           const int32_t* __restrict__ adjwgt,
                                                           const int32_t* __restrict__ part,
                                                           const int32_t* __restrict__ gseg,
                                                           unsigned long long* __restrict__ acc) {
    constexpr int WPB = kBlock / 64;
    __shared__ int32_t s_beg_all[WPB][64], s_part_all[WPB][64];
    __shared__ uint32_t s_cut_all[WPB][64];
    const int wv = threadIdx.x >> 6, lane = __lane_id();
    int32_t* s_beg = s_beg_all[wv];
    int32_t* s_part = s_part_all[wv];
    uint32_t* s_cut = s_cut_all[wv];
    const int64_t ntiles = ((int64_t)n + 63) / 64;
    for (int64_t t = (int64_t)blockIdx.x * WPB + wv; t < ntiles; t += (int64_t)gridDim.x * WPB) {
        const int64_t v0 = t * 64;
        const int last = (int)min<int64_t>(64, (int64_t)n - v0);     // vertices of this tile
        const int64_t v = v0 + (lane < last ? lane : last - 1);
        const int32_t e0 = xadj[v0];
        const int32_t span = xadj[v0 + last] - e0;
        const int32_t gi = gseg[v];
        const bool mixed = __ballot(gi != __shfl(gi, 0, 64)) != 0;
        s_beg[lane] = xadj[v] - e0;
        s_part[lane] = part[v];
        s_cut[lane] = 0u;
        wave_sync();
        uint32_t sum = 0u;
        for (int32_t q = lane; q < span; q += 64) {
            int lo = 0, hi = last - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_beg[mid] <= q) lo = mid;
                else hi = mid - 1;
            }
            const int32_t j = e0 + q;
            const uint32_t w = part[adjncy[j]] != s_part[lo] ? (adjwgt ? (uint32_t)adjwgt[j] : 1u) : 0u;
            if (mixed) {
                if (w) atomicAdd(&s_cut[lo], w);
            } else {
                sum += w;
            }
        }
        if (!mixed) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) sum += (uint32_t)__shfl_xor((int)sum, o, 64);
            if (lane == 0 && sum) atomicAdd(&acc[gi], (unsigned long long)sum);
        } else {
            wave_sync();
            if (lane < last && s_cut[lane]) atomicAdd(&acc[gi], (unsigned long long)s_cut[lane]);
        }
        wave_sync();
    }
}

// The reference sums in i32 with wrap-around and then halves (truncating).
__global__ __launch_bounds__(kBlock) void k_finish_cut_batch(int32_t ng, const unsigned long long* __restrict__ acc,
                                                             int32_t* __restrict__ out) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g < ng) out[g] = (int32_t)(uint32_t)acc[g] / 2;
}

// part[v] = packed[goff[gseg[v]] + v - gv0[gseg[v]]] for the graphs with goff >= 0.
__global__ __launch_bounds__(kBlock) void k_scatter_parts(int32_t n, const int32_t* __restrict__ gseg,
                                                          const int32_t* __restrict__ gv0,
                                                          const int32_t* __restrict__ goff,
                                                          const int32_t* __restrict__ packed,
                                                          int32_t* __restrict__ part) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v >= n) return;
    const int32_t gi = gseg[v];
    const int32_t o = goff[gi];
    if (o >= 0) part[v] = packed[o + (v - gv0[gi])];
}

// One block per listed graph: copies its CSR slice with graph-local ids into
// packed (xadj n+1, adjncy m, vwgt n*ncon at the offsets of its descriptor
// {v0, n, e0, m, out}).
__global__ __launch_bounds__(kBlock) void k_gather_csr(const int32_t* __restrict__ xadj,
                                                       const int32_t* __restrict__ adjncy,
                                                       const int32_t* __restrict__ vwgt, int32_t ncon,
                                                       const int32_t* __restrict__ desc,
                                                       int32_t* __restrict__ packed) {
    const int32_t* d = desc + 5 * blockIdx.x;
    const int32_t v0 = d[0], nn = d[1], e0 = d[2], mm = d[3], o = d[4];
    int32_t* px = packed + o;
    int32_t* pa = px + nn + 1;
    int32_t* pv = pa + mm;
    for (int i = threadIdx.x; i <= nn; i += kBlock) px[i] = xadj[v0 + i] - e0;
    for (int i = threadIdx.x; i < mm; i += kBlock) pa[i] = adjncy[e0 + i] - v0;
    if (vwgt)
        for (int i = threadIdx.x; i < nn * ncon; i += kBlock) pv[i] = vwgt[(int64_t)v0 * ncon + i];
}

// best[v] = part[v] for the vertices of graphs whose attempt improved.
__global__ __launch_bounds__(kBlock) void k_copy_improved(int32_t n, const int32_t* __restrict__ gseg,
                                                          const uint8_t* __restrict__ improved,
                                                          const int32_t* __restrict__ part,
                                                          int32_t* __restrict__ best) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v < n && improved[gseg[v]]) best[v] = part[v];
}

// best[v] = part[a * n + v] for a = choice[gseg[v]] >= 0 (the chosen copy
// of a union replicated once per attempt).
__global__ __launch_bounds__(kBlock) void k_copy_attempt(int32_t n, const int32_t* __restrict__ gseg,
                                                         const int32_t* __restrict__ choice,
                                                         const int32_t* __restrict__ part,
                                                         int32_t* __restrict__ best) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
        const int32_t a = choice[gseg[v]];
        if (a >= 0) best[v] = part[(int64_t)a * n + v];
    }
}

// out[4g..4g+3] = (first vertex, end vertex, first entry, end entry) of
// every graph g with vertices in this level (out cleared beforehand).
__global__ __launch_bounds__(kBlock) void k_graph_ranges(int32_t n, const int32_t* __restrict__ gseg,
                                                         const int32_t* __restrict__ xadj, int32_t* __restrict__ out) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v >= n) return;
    const int32_t gi = gseg[v];
    if (v == 0 || gseg[v - 1] != gi) {
        out[4 * (int64_t)gi] = (int32_t)v;
        out[4 * (int64_t)gi + 2] = xadj[v];
    }
    if (v == n - 1 || gseg[v + 1] != gi) {
        out[4 * (int64_t)gi + 1] = (int32_t)v + 1;
        out[4 * (int64_t)gi + 3] = xadj[v + 1];
    }
}

// One block per graph present: desc {graph, first vertex, n, first tile};
// its tiles of kTile vertices, then padding up to a multiple of 32 tiles.
__global__ __launch_bounds__(kBlock) void k_fill_tiles(const int4* __restrict__ desc, int32_t* __restrict__ tile_v0,
                                                       int32_t* __restrict__ tile_nv, int32_t* __restrict__ tile_g) {
    const int4 d = desc[blockIdx.x];
    const int32_t nt = (d.z + kTile - 1) / kTile;
    const int32_t np = (nt + 31) & ~31;
    for (int i = threadIdx.x; i < np; i += kBlock) {
        const int64_t t = (int64_t)d.w + i;
        const bool real = i < nt;
        tile_v0[t] = real ? d.y + i * kTile : -1;
        tile_nv[t] = real ? min(kTile, d.z - i * kTile) : 0;
        tile_g[t] = real ? d.x : -1;
    }
}

// One block per graph: out[g*ncon + c] = sum of vwgt[v*ncon + c] over its
// vertices [gvtx[g], gvtx[g+1]), as int64.
__global__ __launch_bounds__(kBlock) void k_graph_colsum(const int32_t* __restrict__ gvtx, int32_t ncon,
                                                         const int32_t* __restrict__ vwgt, long long* __restrict__ out,
                                                         long long* __restrict__ outmax) {
    __shared__ long long s_part[kBlock / 64];
    __shared__ int32_t s_max[kBlock / 64];
    const int64_t v0 = gvtx[blockIdx.x], v1 = gvtx[blockIdx.x + 1];
    for (int c = 0; c < ncon; ++c) {
        long long sum = 0;
        int32_t mx = 0;
        for (int64_t v = v0 + threadIdx.x; v < v1; v += kBlock) {
            const int32_t w = vwgt[v * ncon + c];
            sum += w;
            mx = max(mx, w);
        }
        sum = wave_incl_scan(sum);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
        if (__lane_id() == 63) s_part[threadIdx.x >> 6] = sum;
        if (__lane_id() == 0) s_max[threadIdx.x >> 6] = mx;
        __syncthreads();
        if (threadIdx.x == 0) {
            long long t = 0;
            int32_t m = 0;
            for (int w = 0; w < kBlock / 64; ++w) {
                t += s_part[w];
                m = max(m, s_max[w]);
            }
            out[(int64_t)blockIdx.x * ncon + c] = t;
            if (outmax) outmax[(int64_t)blockIdx.x * ncon + c] = m;
        }
        __syncthreads();
    }
}

// One graph's CSR into a union at vertex v0 / entry e0: row offsets shifted
// by e0, neighbour ids by v0.
__global__ __launch_bounds__(kBlock) void k_pack_csr(int32_t n, int64_t m, const int32_t* __restrict__ xadj,
                                                     const int32_t* __restrict__ adjncy, int32_t v0, int32_t e0,
                                                     int32_t* __restrict__ uxadj, int32_t* __restrict__ uadjncy) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) uxadj[v0 + i] = xadj[i] + e0;
    for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < m; j += stride) uadjncy[e0 + j] = adjncy[j] + v0;
}

__global__ __launch_bounds__(kBlock) void k_fill_gseg(int32_t ng, const int32_t* __restrict__ gv0,
                                                      const int32_t* __restrict__ gn, int32_t* __restrict__ gseg) {
    // one block per graph
    const int32_t v0 = gv0[blockIdx.x], nn = gn[blockIdx.x];
    for (int i = threadIdx.x; i < nn; i += kBlock) gseg[v0 + i] = blockIdx.x;
}

__global__ __launch_bounds__(kBlock) void k_column_sum(int32_t n, int32_t ncon, int32_t c,
                                                       const int32_t* __restrict__ vwgt,
                                                       unsigned long long* __restrict__ out,
                                                       unsigned long long* __restrict__ outmax) {
    __shared__ long long s_w[kBlock / 64];
    __shared__ int32_t s_m[kBlock / 64];
    long long sum = 0;
    int32_t mx = 0;
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock) {
        const int32_t w = vwgt[v * ncon + c];
        sum += w;
        mx = max(mx, w);
    }
    long long inc = wave_incl_scan(sum);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
    if (__lane_id() == 63) s_w[threadIdx.x >> 6] = inc;
    if (__lane_id() == 0) s_m[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        int32_t m = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            t += s_w[w];
            m = max(m, s_m[w]);
        }
        atomicAdd(out, (unsigned long long)t);
        if (outmax) atomicMax(outmax, (unsigned long long)m);   // vertex weights are positive
    }
}

__global__ void k_finish_cut(const unsigned long long* acc, int32_t* out) {
    // The reference sums in i32 with wrap-around and then halves (truncating).
    int32_t c = (int32_t)(uint32_t)(*acc);
    *out = c / 2;
}

inline int grid_for(int64_t items, int per_block, int cap = 8192) {
    int64_t b = (items + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

}  // namespace

void edge_cut(const DevGraph& g, const int32_t* part, unsigned long long* acc, int32_t* cut_out,
              int G, hipStream_t s) {
    (void)G;
    RK_HIP(hipMemsetAsync(acc, 0, sizeof(unsigned long long), s));
    const int64_t ntiles = ((int64_t)g.n + 63) / 64;
    const int64_t blocks = (ntiles + kCutWaves - 1) / kCutWaves, resident = (int64_t)device_cus() * kCutBlocksPerCU;
    hipLaunchKernelGGL(k_edge_cut_tile, dim3((unsigned)(blocks < resident ? blocks : resident)),
                       dim3(64 * kCutWaves), 0, s, g.n, g.xadj, g.adjncy, g.adjwgt, part, acc);
    hipLaunchKernelGGL(k_finish_cut, dim3(1), dim3(1), 0, s, acc, cut_out);
    RK_HIP(hipGetLastError());
}

void part_weights(const DevGraph& g, const int32_t* part, int32_t k, int32_t* pw, hipStream_t s, bool zeroed) {
    const int64_t kc = (int64_t)k * g.ncon;
    if (!zeroed) RK_HIP(hipMemsetAsync(pw, 0, kc * sizeof(int32_t), s));
    if (kc <= kPwLds) {
        hipLaunchKernelGGL(k_part_weights_lds, dim3(grid_for(g.n, kPwVerts, 1 << 30)), dim3(kBlock), 0, s,
                           g.n, g.ncon, g.vwgt, part, (int32_t)kc, pw);
    } else {
        hipLaunchKernelGGL(k_part_weights_global, dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, s, g.n,
                           g.ncon, g.vwgt, part, pw);
    }
    RK_HIP(hipGetLastError());
}

void project(int32_t n_fine, const int32_t* cmap, const int32_t* part_c, int32_t* part_f, hipStream_t s,
             int32_t* clear_a, int64_t na, int32_t* clear_b, int32_t nb) {
    if (n_fine <= 0) return;
    hipLaunchKernelGGL(k_project, dim3((n_fine + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n_fine, cmap,
                       part_c, part_f, clear_a, clear_a ? na : 0, clear_b, clear_b ? nb : 0);
    RK_HIP(hipGetLastError());
}

void column_sums(const DevGraph& g, long long* out, hipStream_t s, long long* outmax) {
    RK_HIP(hipMemsetAsync(out, 0, sizeof(long long) * g.ncon, s));
    if (outmax) RK_HIP(hipMemsetAsync(outmax, 0, sizeof(long long) * g.ncon, s));
    for (int c = 0; c < g.ncon; ++c)
        hipLaunchKernelGGL(k_column_sum, dim3(grid_for(g.n, kBlock, 2048)), dim3(kBlock), 0, s, g.n, g.ncon, c,
                           g.vwgt, (unsigned long long*)(out + c),
                           outmax ? (unsigned long long*)(outmax + c) : nullptr);
    RK_HIP(hipGetLastError());
}

void part_weights_batch(const DevGraph& g, const int32_t* part, const int32_t* gseg, const int32_t* gkc,
                        const int32_t* gk, int64_t kc_total, int32_t* pw, hipStream_t s, const uint8_t* gmask,
                        int32_t ngraphs) {
    if (gmask)
        hipLaunchKernelGGL(k_zero_masked, dim3(ngraphs < 1024 ? (ngraphs > 0 ? ngraphs : 1) : 1024), dim3(kBlock), 0,
                           s, ngraphs, g.ncon, gkc, gk, gmask, pw);
    else
        RK_HIP(hipMemsetAsync(pw, 0, kc_total * sizeof(int32_t), s));
    if (g.n > 0)
        hipLaunchKernelGGL(k_part_weights_batch, dim3((g.n + kPwVerts - 1) / kPwVerts), dim3(kBlock), 0, s, g.n,
                           g.ncon, g.vwgt, part, gseg, gkc, gk, gmask, pw);
    RK_HIP(hipGetLastError());
}

void edge_cut_batch(const DevGraph& g, const int32_t* part, const int32_t* gseg, int32_t ngraphs,
                    unsigned long long* acc, int32_t* cut_out, hipStream_t s) {
    RK_HIP(hipMemsetAsync(acc, 0, sizeof(unsigned long long) * ngraphs, s));
    if (g.n > 0)
        hipLaunchKernelGGL(k_edge_cut_batch,
                           dim3((int)std::min<int64_t>(((int64_t)g.n + kBlock - 1) / kBlock, 65536)), dim3(kBlock),
                           0, s, g.n, g.xadj, g.adjncy, g.adjwgt, part, gseg, acc);
    hipLaunchKernelGGL(k_finish_cut_batch, dim3((ngraphs + kBlock - 1) / kBlock), dim3(kBlock), 0, s, ngraphs, acc,
                       cut_out);
    RK_HIP(hipGetLastError());
}

void scatter_parts(int32_t n, const int32_t* gseg, const int32_t* gv0, const int32_t* goff, const int32_t* packed,
                   int32_t* part, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_scatter_parts, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, gseg, gv0, goff,
                           packed, part);
    RK_HIP(hipGetLastError());
}

void gather_csr(const DevGraph& g, const int32_t* desc, int32_t ndesc, int32_t* packed, hipStream_t s) {
    if (ndesc > 0)
        hipLaunchKernelGGL(k_gather_csr, dim3(ndesc), dim3(kBlock), 0, s, g.xadj, g.adjncy, g.vwgt, g.ncon, desc,
                           packed);
    RK_HIP(hipGetLastError());
}

void copy_improved(int32_t n, const int32_t* gseg, const uint8_t* improved, const int32_t* part, int32_t* best,
                   hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_copy_improved, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, gseg, improved,
                           part, best);
    RK_HIP(hipGetLastError());
}

void copy_attempt(int32_t n, const int32_t* gseg, const int32_t* choice, const int32_t* part, int32_t* best,
                  hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_copy_attempt, dim3((unsigned)std::min<int64_t>(((int64_t)n + kBlock - 1) / kBlock, 16384)),
                           dim3(kBlock), 0, s, n, gseg, choice, part, best);
    RK_HIP(hipGetLastError());
}

void graph_ranges(int32_t n, const int32_t* gseg, const int32_t* xadj, int32_t ngraphs_total, int32_t* out,
                  hipStream_t s) {
    RK_HIP(hipMemsetAsync(out, 0, sizeof(int32_t) * 4 * (size_t)ngraphs_total, s));
    if (n > 0)
        hipLaunchKernelGGL(k_graph_ranges, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, gseg, xadj, out);
    RK_HIP(hipGetLastError());
}

void fill_tiles(const int4* desc, int32_t ndesc, int32_t* tile_v0, int32_t* tile_nv, int32_t* tile_g,
                hipStream_t s) {
    if (ndesc > 0)
        hipLaunchKernelGGL(k_fill_tiles, dim3(ndesc), dim3(kBlock), 0, s, desc, tile_v0, tile_nv, tile_g);
    RK_HIP(hipGetLastError());
}

void graph_column_sums(const DevGraph& g, const int32_t* gvtx, int32_t ngraphs, long long* out, hipStream_t s,
                       long long* outmax) {
    if (ngraphs > 0)
        hipLaunchKernelGGL(k_graph_colsum, dim3(ngraphs), dim3(kBlock), 0, s, gvtx, g.ncon, g.vwgt, out, outmax);
    RK_HIP(hipGetLastError());
}

void pack_csr(int32_t n, int64_t m, const int32_t* xadj, const int32_t* adjncy, int32_t v0, int32_t e0,
              int32_t* uxadj, int32_t* uadjncy, hipStream_t s) {
    const int64_t work = std::max<int64_t>(n, m);
    if (work > 0)
        hipLaunchKernelGGL(k_pack_csr, dim3((unsigned)std::min<int64_t>((work + kBlock - 1) / kBlock, 8192)),
                           dim3(kBlock), 0, s, n, m, xadj, adjncy, v0,
                           e0, uxadj, uadjncy);
    RK_HIP(hipGetLastError());
}

void fill_gseg(int32_t ngraphs, const int32_t* gv0, const int32_t* gn, int32_t* gseg, hipStream_t s) {
    if (ngraphs > 0)
        hipLaunchKernelGGL(k_fill_gseg, dim3(ngraphs), dim3(kBlock), 0, s, ngraphs, gv0, gn, gseg);
    RK_HIP(hipGetLastError());
}

int pick_group(int32_t n, int32_t m) {
    double avg = n > 0 ? (double)m / (double)n : 0.0;
    if (avg <= 4.5) return 4;
    if (avg <= 10.0) return 8;
    if (avg <= 24.0) return 16;
    if (avg <= 48.0) return 32;
    return 64;
}

bool validate_csr_device(const DevGraph& g, int32_t* cnt, int32_t* bad_dev, int32_t* bad_host, hipStream_t s) {
    RK_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * (size_t)g.n, s));
    RK_HIP(hipMemsetAsync(bad_dev, 0, sizeof(int32_t), s));
    const int grid = (int)std::min<int64_t>(((int64_t)g.n + kBlock - 1) / kBlock, 8192);
    hipLaunchKernelGGL(k_validate_rows, dim3(grid), dim3(kBlock), 0, s, g.n, (int64_t)g.m, g.xadj, g.adjncy, cnt,
                       bad_dev);
    hipLaunchKernelGGL(k_validate_counts, dim3(grid), dim3(kBlock), 0, s, g.n, g.ncon, g.xadj, g.vwgt, cnt, bad_dev);
    RK_HIP(hipGetLastError());
    RK_HIP(hipMemcpyAsync(bad_host, bad_dev, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RK_HIP(hipStreamSynchronize(s));
    return *bad_host == 0;
}

}  // namespace rk
