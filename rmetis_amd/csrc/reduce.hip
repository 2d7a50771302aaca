// reduce.hip -- edge-cut and part-weight reductions, partition projection.
//
// edge cut:      graph/mod.rs:123-137 (sum over directed entries, i32 wrap, /2)
// part weights:  graph/mod.rs:140-149 (pw[part[v]][c] += vwgt[v][c], i32 wrap)
// projection:    partition/kway.rs:81-84 (fine_part[v] = part[cmap[v]])
#include "common.hpp"
#include "kernels.hpp"

namespace rk {

namespace {

// G lanes per vertex, blocks stride over vertices.  Each lane keeps an int64
// partial; the block sum is added to one 64-bit accumulator (integer addition
// is order independent, so the result is deterministic).
template <int G>
__global__ __launch_bounds__(kBlock) void k_edge_cut(int32_t n, const int32_t* __restrict__ xadj,
                                                     const int32_t* __restrict__ adjncy,
                                                     const int32_t* __restrict__ adjwgt,
                                                     const int32_t* __restrict__ part,
                                                     unsigned long long* __restrict__ acc) {
    constexpr int NG = kBlock / G;
    __shared__ long long s_w[kBlock / 64];
    const int gl = group_lane<G>();
    const int grp = threadIdx.x / G;
    long long sum = 0;
    for (int64_t v = (int64_t)blockIdx.x * NG + grp; v < n; v += (int64_t)gridDim.x * NG) {
        const int32_t pv = part[v];
        const int32_t beg = xadj[v], end = xadj[v + 1];
        for (int32_t j = beg + gl; j < end; j += G) {
            if (part[adjncy[j]] != pv) sum += adjwgt ? adjwgt[j] : 1;
        }
    }
    long long inc = wave_incl_scan(sum);
    if (__lane_id() == 63) s_w[threadIdx.x >> 6] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += s_w[w];
        atomicAdd(acc, (unsigned long long)t);
    }
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

__global__ __launch_bounds__(kBlock) void k_project(int32_t n, const int32_t* __restrict__ cmap,
                                                    const int32_t* __restrict__ part_c,
                                                    int32_t* __restrict__ part_f) {
    int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v < n) part_f[v] = part_c[cmap[v]];
}

__global__ __launch_bounds__(kBlock) void k_column_sum(int32_t n, int32_t ncon, int32_t c,
                                                       const int32_t* __restrict__ vwgt,
                                                       unsigned long long* __restrict__ out) {
    __shared__ long long s_w[kBlock / 64];
    long long sum = 0;
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock)
        sum += vwgt[v * ncon + c];
    long long inc = wave_incl_scan(sum);
    if (__lane_id() == 63) s_w[threadIdx.x >> 6] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += s_w[w];
        atomicAdd(out, (unsigned long long)t);
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
    RK_HIP(hipMemsetAsync(acc, 0, sizeof(unsigned long long), s));
    const int grid = grid_for(g.n, kBlock / G, 16384);
    switch (G) {
        case 4: hipLaunchKernelGGL(k_edge_cut<4>, dim3(grid), dim3(kBlock), 0, s, g.n, g.xadj, g.adjncy, g.adjwgt, part, acc); break;
        case 8: hipLaunchKernelGGL(k_edge_cut<8>, dim3(grid), dim3(kBlock), 0, s, g.n, g.xadj, g.adjncy, g.adjwgt, part, acc); break;
        case 16: hipLaunchKernelGGL(k_edge_cut<16>, dim3(grid), dim3(kBlock), 0, s, g.n, g.xadj, g.adjncy, g.adjwgt, part, acc); break;
        case 32: hipLaunchKernelGGL(k_edge_cut<32>, dim3(grid), dim3(kBlock), 0, s, g.n, g.xadj, g.adjncy, g.adjwgt, part, acc); break;
        default: hipLaunchKernelGGL(k_edge_cut<64>, dim3(grid), dim3(kBlock), 0, s, g.n, g.xadj, g.adjncy, g.adjwgt, part, acc); break;
    }
    hipLaunchKernelGGL(k_finish_cut, dim3(1), dim3(1), 0, s, acc, cut_out);
    RK_HIP(hipGetLastError());
}

void part_weights(const DevGraph& g, const int32_t* part, int32_t k, int32_t* pw, hipStream_t s) {
    const int64_t kc = (int64_t)k * g.ncon;
    RK_HIP(hipMemsetAsync(pw, 0, kc * sizeof(int32_t), s));
    if (kc <= kPwLds) {
        hipLaunchKernelGGL(k_part_weights_lds, dim3(grid_for(g.n, kPwVerts, 1 << 30)), dim3(kBlock), 0, s,
                           g.n, g.ncon, g.vwgt, part, (int32_t)kc, pw);
    } else {
        hipLaunchKernelGGL(k_part_weights_global, dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, s, g.n,
                           g.ncon, g.vwgt, part, pw);
    }
    RK_HIP(hipGetLastError());
}

void project(int32_t n_fine, const int32_t* cmap, const int32_t* part_c, int32_t* part_f, hipStream_t s) {
    if (n_fine <= 0) return;
    hipLaunchKernelGGL(k_project, dim3((n_fine + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n_fine, cmap,
                       part_c, part_f);
    RK_HIP(hipGetLastError());
}

void column_sums(const DevGraph& g, long long* out, hipStream_t s) {
    RK_HIP(hipMemsetAsync(out, 0, sizeof(long long) * g.ncon, s));
    for (int c = 0; c < g.ncon; ++c)
        hipLaunchKernelGGL(k_column_sum, dim3(grid_for(g.n, kBlock, 2048)), dim3(kBlock), 0, s, g.n, g.ncon, c,
                           g.vwgt, (unsigned long long*)(out + c));
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

}  // namespace rk
