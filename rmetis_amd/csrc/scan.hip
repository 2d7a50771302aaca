// scan.hip -- device-wide exclusive prefix sum of int32 (reduce-then-scan).
//
// Used for the contraction's coarse numbering (ascending-leader order of
// graph/mod.rs:288-297), coarse CSR offsets and the ordered compaction of
// greedy move candidates.  Up to kScanOneMax elements one workgroup of 1024
// threads scans the input in one launch (the coarser levels are small and
// launch-latency bound).  Larger
// inputs take three launches: per-chunk sums, a single-block scan of the
// chunk sums, and a per-chunk rescan that adds the chunk offset.
#include "common.hpp"
#include "kernels.hpp"

namespace rk {

namespace {

constexpr int kScanItems = 16;
constexpr int kScanChunk = kBlock * kScanItems;   // 4096 elements per block
constexpr int kChunkThreads = 1024;               // the chunk-sum scan: one block ...
constexpr int kChunkItems = 16;                   // ... of 16 consecutive sums per thread and round

constexpr int kOneThreads = 1024;
constexpr int kOneRows = 16;
constexpr int64_t kScanOneMax = (int64_t)kOneThreads * kOneRows;   // 16,384

// Exclusive scan over the NT threads of the block; *total = the block's sum.
// Both barriers are inside, so consecutive calls may reuse s_w.
template <int NT = kBlock>
__device__ inline long long block_excl_scan_ll(long long x, long long* s_w, long long* total) {
    const int lane = __lane_id();
    const int wave = threadIdx.x >> 6;
    long long inc = wave_incl_scan(x);
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    long long off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) {
        long long c = s_w[w];
        if (w < wave) off += c;
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return off + inc - x;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const int32_t* __restrict__ in, int64_t n,
                                                        long long* __restrict__ chunk_sums) {
    __shared__ long long s_w[kBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanChunk;
    long long sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t idx = base + (int64_t)i * kBlock + threadIdx.x;
        if (idx < n) sum += in[idx];
    }
    long long tot;
    block_excl_scan_ll(sum, s_w, &tot);
    if (threadIdx.x == 0) chunk_sums[blockIdx.x] = tot;
}

// The chunk sums in place to their exclusive prefix: rounds of
// kChunkThreads * kChunkItems sums, each thread scanning kChunkItems
// consecutive ones between one block scan of the threads' totals.
__global__ __launch_bounds__(kChunkThreads) void k_scan_chunks(long long* __restrict__ chunk_sums, int nchunks,
                                                               int32_t* __restrict__ total_out) {
    __shared__ long long s_w[kChunkThreads / 64];
    long long carry = 0;
    for (int b = 0; b < nchunks; b += kChunkThreads * kChunkItems) {
        const int i0 = b + threadIdx.x * kChunkItems;
        long long x[kChunkItems];
        long long sum = 0;
#pragma unroll
        for (int i = 0; i < kChunkItems; ++i) {
            x[i] = i0 + i < nchunks ? chunk_sums[i0 + i] : 0;
            sum += x[i];
        }
        long long tot;
        long long run = carry + block_excl_scan_ll<kChunkThreads>(sum, s_w, &tot);
#pragma unroll
        for (int i = 0; i < kChunkItems; ++i) {
            if (i0 + i < nchunks) chunk_sums[i0 + i] = run;
            run += x[i];
        }
        carry += tot;
    }
    if (threadIdx.x == 0 && total_out) *total_out = (int32_t)carry;
}

// Each thread rescans kScanItems consecutive elements of its block's chunk;
// VEC: in and out are 16-byte aligned, and a thread's whole run inside the
// input moves as 16-byte vectors.
template <bool VEC>
__global__ __launch_bounds__(kBlock) void k_scan_down(const int32_t* __restrict__ in, int64_t n,
                                                      const long long* __restrict__ chunk_off,
                                                      int32_t* __restrict__ out) {
    __shared__ long long s_w[kBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanItems;
    const bool whole = VEC && base + kScanItems <= n;
    int32_t v[kScanItems];
    if (whole) {
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            const int4 t = *reinterpret_cast<const int4*>(in + base + 4 * q);
            v[4 * q] = t.x, v[4 * q + 1] = t.y, v[4 * q + 2] = t.z, v[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) v[i] = base + i < n ? in[base + i] : 0;
    }
    long long sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) sum += v[i];
    long long tot;
    long long run = block_excl_scan_ll(sum, s_w, &tot) + chunk_off[blockIdx.x];
    int32_t o[kScanItems];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        o[i] = (int32_t)run;
        run += v[i];
    }
    if (whole) {
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q)
            *reinterpret_cast<int4*>(out + base + 4 * q) = make_int4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; ++i)
            if (base + i < n) out[base + i] = o[i];
    }
}

// n <= kScanOneMax, one block of kOneThreads, one round.  The input is read
// as rows of kOneThreads elements (thread t loads element j*kOneThreads + t
// of every row j, so each load is coalesced); each wave scans its 64-element
// piece of every row, and one block scan over the pieces' totals in (row,
// wave) order -- which is element order -- gives each piece's offset.  Sums
// are taken modulo 2^32: the outputs are int32 truncations of the exact
// prefix sums, so the low 32 bits are all that matter.  Every load precedes
// every store, so in == out is allowed.
__global__ __launch_bounds__(kOneThreads) void k_scan_one(const int32_t* in, int n, int32_t* out,
                                                          int32_t* total_out) {
    __shared__ uint32_t s_seg[kOneThreads];   // piece (row j, wave w) at j * 16 + w
    __shared__ long long s_w[kOneThreads / 64];
    const int t = threadIdx.x, lane = __lane_id(), wave = t >> 6;
    const int rows = (n + kOneThreads - 1) / kOneThreads;   // <= kOneRows
    uint32_t x[kOneRows];
#pragma unroll
    for (int j = 0; j < kOneRows; ++j) {
        const int idx = j * kOneThreads + t;
        x[j] = j < rows && idx < n ? (uint32_t)in[idx] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kOneRows; ++j) {
        if (j < rows) {
            uint32_t inc = x[j];
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(inc, o, 64);
                if (lane >= o) inc += y;
            }
            if (lane == 63) s_seg[j * (kOneThreads / 64) + wave] = inc;
            x[j] = inc - x[j];   // exclusive within the piece
        }
    }
    __syncthreads();
    const uint32_t piece = t < rows * (kOneThreads / 64) ? s_seg[t] : 0u;
    long long tot;
    const long long off = block_excl_scan_ll<kOneThreads>((long long)piece, s_w, &tot);
    s_seg[t] = (uint32_t)off;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kOneRows; ++j) {
        const int idx = j * kOneThreads + t;
        if (j < rows && idx < n) out[idx] = (int32_t)(s_seg[j * (kOneThreads / 64) + wave] + x[j]);
    }
    if (t == 0 && total_out) *total_out = (int32_t)tot;
}

}  // namespace

size_t scan_temp_bytes(int64_t n) {
    int64_t nchunks = (n + kScanChunk - 1) / kScanChunk;
    return (size_t)(nchunks + 1) * sizeof(long long);
}

void exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, int32_t* total_dev, void* temp,
                        hipStream_t s) {
    if (n <= 0) {
        if (total_dev) RK_HIP(hipMemsetAsync(total_dev, 0, sizeof(int32_t), s));
        return;
    }
    if (n <= kScanOneMax) {
        hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(kOneThreads), 0, s, in, (int)n, out, total_dev);
        RK_HIP(hipGetLastError());
        return;
    }
    int nchunks = (int)((n + kScanChunk - 1) / kScanChunk);
    long long* sums = (long long*)temp;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nchunks), dim3(kBlock), 0, s, in, n, sums);
    hipLaunchKernelGGL(k_scan_chunks, dim3(1), dim3(kChunkThreads), 0, s, sums, nchunks, total_dev);
    if (((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0)
        hipLaunchKernelGGL(k_scan_down<true>, dim3(nchunks), dim3(kBlock), 0, s, in, n, sums, out);
    else
        hipLaunchKernelGGL(k_scan_down<false>, dim3(nchunks), dim3(kBlock), 0, s, in, n, sums, out);
    RK_HIP(hipGetLastError());
}

}  // namespace rk
