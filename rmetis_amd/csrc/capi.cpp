// capi.cpp -- extern "C" boundary of librmetis_hip.so (include/rmetis.h).
//
// RMETIS_SetDefaultOptions / RMETIS_PartGraphKway follow src/ffi/c_api.rs:
// NULL checks and n/ncon/nparts checks (157-167), optional arrays as None
// (169-177), option parsing (211-222), CSR validation with
// graph/validate.rs:16-142 (186-187), part/objval written only on success
// (196-204) and the error-code mapping (205-207).  Every call runs on its own
// pooled stream and arena, so concurrent calls are safe.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rmetis.h"
#include "driver.hpp"
#include "graphio.hpp"
#include "tuning.hpp"
#include "profiler.hpp"
#include "serve.hpp"

namespace rk {

namespace {
std::mutex g_pool_mu;
std::vector<Context*> g_pool;
thread_local std::string g_last_error;
}  // namespace

Context* acquire_context() {
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt <= 0)
        throw Error(-4, "no HIP device is available (the k-way path runs only on the GPU)");
    int dev = 0;
    RK_HIP(hipGetDevice(&dev));
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        for (size_t i = 0; i < g_pool.size(); ++i)
            if (g_pool[i]->device == dev) {
                Context* c = g_pool[i];
                g_pool.erase(g_pool.begin() + i);
                return c;
            }
    }
    Context* c = new Context();
    c->device = dev;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        throw Error(-4, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
    }
    return c;
}

void release_context(Context* c) {
    c->phase = nullptr;   // an observer's counter never outlives the call that set it
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool.push_back(c);
}

namespace {

template <typename F>
int32_t guarded(F&& f) {
    try {
        g_last_error.clear();
        return f();
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host out of memory";
        return -3;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return -4;
    } catch (...) {
        g_last_error = "unknown error";
        return -4;
    }
}

// graph/validate.rs:16-142.  Returns false for MetisError::Input.
// Runs body(t, lo, hi) over [0, count) in chunks on up to 16 host threads
// (one for small inputs); returns true when every chunk returned true.
template <typename F>
bool parallel_all(int64_t count, int64_t work, F body) {
    int nt = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (work < (1 << 20)) nt = 1;
    if (nt == 1) return body(0, (int64_t)0, count);
    std::vector<char> ok(nt, 1);
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; ++t) {
        const int64_t lo = count * t / nt, hi = count * (t + 1) / nt;
        th.emplace_back([&, t, lo, hi]() { ok[t] = body(t, lo, hi) ? 1 : 0; });
    }
    for (auto& x : th) x.join();
    for (char c : ok)
        if (!c) return false;
    return true;
}

// graph/validate.rs:16-142: xadj starts at 0 and never decreases, every
// neighbour is in range and not the vertex itself, every vertex appears in
// as many rows as its own row is long (the count check of a symmetric
// adjacency), vertex weights are positive.  Large graphs are checked on
// several host threads.
bool validate_csr(int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy, const int32_t* vwgt) {
    if (n <= 0 || ncon <= 0) return false;
    if (xadj[0] != 0) return false;
    const int64_t work = (int64_t)n + std::max<int64_t>(0, xadj[n]);
    if (!parallel_all(n, work, [&](int, int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; ++i)
                if (xadj[i] > xadj[i + 1]) return false;
            return true;
        }))
        return false;
    // per-vertex in-degree counts, incremented by every thread
    std::unique_ptr<std::atomic<int32_t>[]> cnt(new std::atomic<int32_t>[n]);
    if (!parallel_all(n, work, [&](int, int64_t lo, int64_t hi) {
            for (int64_t v = lo; v < hi; ++v) cnt[v].store(0, std::memory_order_relaxed);
            return true;
        }))
        return false;
    if (!parallel_all(n, work, [&](int, int64_t lo, int64_t hi) {
            for (int64_t v = lo; v < hi; ++v)
                for (int32_t j = xadj[v]; j < xadj[v + 1]; ++j) {
                    const int32_t u = adjncy[j];
                    if (u < 0 || u >= n || u == (int32_t)v) return false;
                    cnt[u].fetch_add(1, std::memory_order_relaxed);
                }
            return true;
        }))
        return false;
    if (!parallel_all(n, work, [&](int, int64_t lo, int64_t hi) {
            for (int64_t v = lo; v < hi; ++v)
                if (cnt[v].load(std::memory_order_relaxed) != xadj[v + 1] - xadj[v]) return false;
            return true;
        }))
        return false;
    if (vwgt &&
        !parallel_all((int64_t)n * ncon, work, [&](int, int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; ++i)
                if (vwgt[i] <= 0) return false;
            return true;
        }))
        return false;
    return true;
}

// xadj[0] == 0 and xadj never decreases (the first checks of
// graph/validate.rs), on several host threads for large graphs.
bool xadj_valid(int32_t n, const int32_t* xadj) {
    if (n <= 0 || xadj[0] != 0) return false;
    return parallel_all(n, n, [&](int, int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i)
            if (xadj[i] > xadj[i + 1]) return false;
        return true;
    });
}

// Per-constraint vertex-weight totals below 2^30, so that running part
// weights and weight+vw never leave the int32 range in the greedy kernels.
bool totals_in_range(int32_t n, int32_t ncon, const int32_t* vwgt, std::vector<int64_t>* tot = nullptr) {
    std::vector<int64_t> t(ncon, 0);
    for (int64_t v = 0; v < n; ++v)
        for (int c = 0; c < ncon; ++c) t[c] += vwgt ? vwgt[v * ncon + c] : 1;
    for (int c = 0; c < ncon; ++c)
        if (t[c] >= (1ll << 30)) return false;
    if (tot) *tot = t;
    return true;
}

bool valid_part(int32_t n, const int32_t* part, int32_t nparts) {
    for (int32_t v = 0; v < n; ++v)
        if (part[v] < 0 || part[v] >= nparts) return false;
    return true;
}

template <typename T>
T* upload(Context& cx, const T* h, int64_t count) {
    if (!h) return nullptr;
    T* d = cx.arena.get<T>(count);
    if (count > 0) RK_HIP(hipMemcpyAsync(d, h, sizeof(T) * count, hipMemcpyHostToDevice, cx.stream));
    return d;
}

template <typename T>
void download(Context& cx, T* h, const T* d, int64_t count) {
    if (count > 0) RK_HIP(hipMemcpyAsync(h, d, sizeof(T) * count, hipMemcpyDeviceToHost, cx.stream));
}

DevGraph upload_graph(Context& cx, int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                      const int32_t* vwgt, const int32_t* adjwgt, const int32_t* vsize) {
    DevGraph g;
    g.n = n;
    g.ncon = ncon;
    g.m = xadj[n];
    g.xadj = upload(cx, xadj, (int64_t)n + 1);
    g.adjncy = upload(cx, adjncy, g.m);
    g.vwgt = upload(cx, vwgt, (int64_t)n * ncon);
    g.adjwgt = upload(cx, adjwgt, g.m);
    g.vsize = upload(cx, vsize, n);
    return g;
}

void fill_stats(const KwayStats& ks, rmetis_hip_stats* out) {
    if (!out) return;
    std::memset(out, 0, sizeof(*out));
    out->nlevels = (int32_t)std::min<size_t>(ks.levels.size(), RMETIS_HIP_MAX_LEVELS);
    out->coarsest_n = ks.coarsest_n;
    out->initial_ms = ks.initial_ms;
    out->attempts_per_call = ks.attempts_per_call;
    for (int i = 0; i < out->nlevels; ++i) {
        out->level_n[i] = ks.levels[i].n;
        out->level_m[i] = ks.levels[i].m;
        out->coarsen_ms[i] = ks.levels[i].coarsen_ms;
        out->refine_ms[i] = ks.levels[i].refine_ms;
        out->passes[i] = ks.levels[i].passes;
    }
}

// Fixed device workspace of the allocation-free k-way entry point: one
// array per GPU (each device holds its own copy from the library's load to
// the end of the process).  Host code never uses the array's name as an
// address: its device address is looked up once per device with
// hipGetSymbolAddress and kept in g_scratch_ptr.
constexpr size_t kScratchBytes = 128ull << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu[kMaxDevices];      // one call at a time per device
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on device dev, the current device (caller
// holds g_scratch_mu[dev]).
void* scratch_address(int dev) {
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(-4, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace rk

using namespace rk;

struct rmetis_hip_server {
    std::unique_ptr<rk::Server> s;
    int32_t device = 0;
};

extern "C" {

int32_t RMETIS_SetDefaultOptions(idx_t* options) {
    if (!options) return METIS_ERROR_INPUT;
    default_options_raw(options);
    return METIS_OK;
}

int32_t RMETIS_PartGraphKway(const idx_t* nvtxs, const idx_t* ncon, const idx_t* xadj, const idx_t* adjncy,
                             const idx_t* vwgt, const idx_t* vsize, const idx_t* adjwgt, const idx_t* nparts,
                             const real_t* tpwgts, const real_t* ubvec, const idx_t* options, idx_t* objval,
                             idx_t* part) {
    return guarded([&]() -> int32_t {
        if (!nvtxs || !ncon || !xadj || !adjncy || !nparts || !part) return METIS_ERROR_INPUT;
        const int32_t n = *nvtxs, nc = *ncon, np = *nparts;
        if (n <= 0 || nc <= 0 || np < 2) return METIS_ERROR_INPUT;
        // validate.rs:16-142: xadj on the host (it sizes the upload), the
        // rows and weights on the device once uploaded
        if (!xadj_valid(n, xadj)) return METIS_ERROR_INPUT;
        if (np > n) return METIS_ERROR_INPUT;                       // kway.rs:34-38 (either way an input error)
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
            // no device to run on: still report a malformed graph as such
            if (!validate_csr(n, nc, xadj, adjncy, vwgt)) return METIS_ERROR_INPUT;
        }
        const Options o = parse_options(options);
        ContextGuard cx;
        DevGraph g = upload_graph(*cx.c, n, nc, xadj, adjncy, vwgt, adjwgt, nullptr);
        if (!validate_csr_device(g, cx->arena.get<int32_t>(n), cx->arena.get<int32_t>(1), cx->host_small(),
                                 cx->stream))
            return METIS_ERROR_INPUT;
        (void)vsize;  // the k-way greedy path never reads vsize
        int32_t* d_part = cx->arena.get<int32_t>(n);
        const int32_t obj = kway_device(*cx.c, g, np, tpwgts, ubvec, o, d_part, nullptr);
        std::vector<int32_t> hp(n);
        download(*cx.c, hp.data(), d_part, n);
        RK_HIP(hipStreamSynchronize(cx->stream));
        std::copy(hp.begin(), hp.end(), part);
        if (objval) *objval = obj;
        return METIS_OK;
    });
}

int32_t RMETIS_PartGraphRecursive(const idx_t* nvtxs, const idx_t* ncon, const idx_t* xadj, const idx_t* adjncy,
                                  const idx_t* vwgt, const idx_t* vsize, const idx_t* adjwgt, const idx_t* nparts,
                                  const real_t* tpwgts, const real_t* ubvec, const idx_t* options, idx_t* objval,
                                  idx_t* part) {
    // c_api.rs:25-48 -> c_partition_impl(kway = false) -> partition_recursive
    return guarded([&]() -> int32_t {
        if (!nvtxs || !ncon || !xadj || !adjncy || !nparts || !part) return METIS_ERROR_INPUT;   // c_api.rs:157-161
        const int32_t n = *nvtxs, nc = *ncon, np = *nparts;
        if (n <= 0 || nc <= 0 || np < 2) return METIS_ERROR_INPUT;                            // c_api.rs:167
        if (!validate_csr(n, nc, xadj, adjncy, vwgt)) return METIS_ERROR_INPUT;              // c_api.rs:186-187
        if (np > n) return METIS_ERROR_INPUT;                                                 // recursive.rs:22-26
        if (!totals_in_range(n, nc, vwgt)) return METIS_ERROR_INPUT;
        const Options o = parse_options(options);
        ContextGuard cx;
        DevGraph g = upload_graph(*cx.c, n, nc, xadj, adjncy, vwgt, adjwgt, nullptr);
        (void)vsize;  // the edge-cut objective never reads vsize
        int32_t* d_part = cx->arena.get<int32_t>(n);
        const int32_t obj = recursive_device(*cx.c, g, np, tpwgts, ubvec, o, d_part);
        std::vector<int32_t> hp(n);
        download(*cx.c, hp.data(), d_part, n);
        RK_HIP(hipStreamSynchronize(cx->stream));
        std::copy(hp.begin(), hp.end(), part);
        if (objval) *objval = obj;
        return METIS_OK;
    });
}

int32_t RMETIS_NodeND(const idx_t* nvtxs, const idx_t* xadj, const idx_t* adjncy, const idx_t* vwgt,
                      const idx_t* options, idx_t* perm, idx_t* iperm) {
    // c_api.rs:83-134 -> nd::node_nd (partition/nd.rs:16-44)
    return guarded([&]() -> int32_t {
        if (!nvtxs || !xadj || !adjncy || !perm || !iperm) return METIS_ERROR_INPUT;
        const int32_t n = *nvtxs;
        if (n <= 0) return METIS_ERROR_INPUT;
        if (!validate_csr(n, 1, xadj, adjncy, vwgt)) return METIS_ERROR_INPUT;   // Graph::new(n, 1, ...)
        if (!totals_in_range(n, 1, vwgt)) return METIS_ERROR_INPUT;
        const Options o = parse_options(options);
        ContextGuard cx;
        DevGraph g = upload_graph(*cx.c, n, 1, xadj, adjncy, vwgt, nullptr, nullptr);
        int32_t* d_perm = cx->arena.get<int32_t>(n);
        int32_t* d_iperm = cx->arena.get<int32_t>(n);
        node_nd_device(*cx.c, g, o, d_perm, d_iperm);
        std::vector<int32_t> hp(n), hi(n);
        download(*cx.c, hp.data(), d_perm, n);
        download(*cx.c, hi.data(), d_iperm, n);
        RK_HIP(hipStreamSynchronize(cx->stream));
        std::copy(hp.begin(), hp.end(), perm);
        std::copy(hi.begin(), hi.end(), iperm);
        return METIS_OK;
    });
}

int32_t rmetis_hip_recursive_device(idx_t nvtxs, idx_t ncon, const idx_t* d_xadj, const idx_t* d_adjncy,
                                    const idx_t* d_vwgt, const idx_t* d_adjwgt, idx_t nparts, const real_t* tpwgts,
                                    const real_t* ubvec, const idx_t* options, idx_t* objval, idx_t* d_part) {
    return guarded([&]() -> int32_t {
        if (!d_xadj || !d_adjncy || !d_part) return METIS_ERROR_INPUT;
        if (nvtxs <= 0 || ncon <= 0 || nparts < 2) return METIS_ERROR_INPUT;
        const Options o = parse_options(options);
        ContextGuard cx;
        DevGraph g;
        g.n = nvtxs;
        g.ncon = ncon;
        RK_HIP(hipMemcpyAsync(&g.m, d_xadj + nvtxs, sizeof(int32_t), hipMemcpyDeviceToHost, cx->stream));
        RK_HIP(hipStreamSynchronize(cx->stream));
        g.xadj = (int32_t*)d_xadj;
        g.adjncy = (int32_t*)d_adjncy;
        g.vwgt = (int32_t*)d_vwgt;
        g.adjwgt = (int32_t*)d_adjwgt;
        const int32_t obj = recursive_device(*cx.c, g, nparts, tpwgts, ubvec, o, d_part);
        if (objval) *objval = obj;
        return METIS_OK;
    });
}

int32_t rmetis_hip_nodend_device(idx_t nvtxs, const idx_t* d_xadj, const idx_t* d_adjncy, const idx_t* d_vwgt,
                                 const idx_t* options, idx_t* d_perm, idx_t* d_iperm) {
    return guarded([&]() -> int32_t {
        if (!d_xadj || !d_adjncy || !d_perm || !d_iperm || nvtxs <= 0) return METIS_ERROR_INPUT;
        const Options o = parse_options(options);
        ContextGuard cx;
        DevGraph g;
        g.n = nvtxs;
        g.ncon = 1;
        RK_HIP(hipMemcpyAsync(&g.m, d_xadj + nvtxs, sizeof(int32_t), hipMemcpyDeviceToHost, cx->stream));
        RK_HIP(hipStreamSynchronize(cx->stream));
        g.xadj = (int32_t*)d_xadj;
        g.adjncy = (int32_t*)d_adjncy;
        g.vwgt = (int32_t*)d_vwgt;
        node_nd_device(*cx.c, g, o, d_perm, d_iperm);
        return METIS_OK;
    });
}

int32_t rmetis_read_graph(const char* path, idx_t* nvtxs, idx_t* ncon, idx_t** xadj, idx_t** adjncy, idx_t** vwgt,
                          idx_t** vsize, idx_t** adjwgt) {
    return guarded([&]() -> int32_t {
        if (!path || !nvtxs || !ncon || !xadj || !adjncy) return METIS_ERROR_INPUT;
        const GraphFile f = read_graph_file(path);
        auto out = [](const std::vector<int32_t>& v, bool has) -> idx_t* {
            if (!has) return nullptr;
            idx_t* p = (idx_t*)std::malloc(sizeof(idx_t) * std::max<size_t>(v.size(), 1));
            if (!p) throw std::bad_alloc();
            std::copy(v.begin(), v.end(), p);
            return p;
        };
        idx_t* x = out(f.xadj, true);
        idx_t* a = out(f.adjncy, true);
        idx_t* w = out(f.vwgt, f.has_vwgt);
        idx_t* z = out(f.vsize, f.has_vsize);
        idx_t* e = out(f.adjwgt, f.has_adjwgt);
        *nvtxs = f.n;
        *ncon = f.ncon;
        *xadj = x;
        *adjncy = a;
        if (vwgt) *vwgt = w;
        else std::free(w);
        if (vsize) *vsize = z;
        else std::free(z);
        if (adjwgt) *adjwgt = e;
        else std::free(e);
        return METIS_OK;
    });
}

int32_t rmetis_write_graph(const char* path, idx_t nvtxs, idx_t ncon, const idx_t* xadj, const idx_t* adjncy,
                           const idx_t* vwgt, const idx_t* vsize, const idx_t* adjwgt) {
    return guarded([&]() -> int32_t {
        if (!path) return METIS_ERROR_INPUT;
        write_graph_file(path, nvtxs, ncon, xadj, adjncy, vwgt, vsize, adjwgt);
        return METIS_OK;
    });
}

void rmetis_free(void* p) { std::free(p); }

int32_t rmetis_hip_fm2_refine(idx_t nvtxs, idx_t ncon, const idx_t* xadj, const idx_t* adjncy, const idx_t* vwgt,
                              const idx_t* adjwgt, idx_t* part, const double* tpwgts, const double* ubvec, idx_t niter,
                              idx_t* cut) {
    // fm_refine, src/refine/fm.rs:12-29
    return guarded([&]() -> int32_t {
        if (!xadj || !adjncy || !part || !tpwgts || !ubvec || nvtxs <= 0 || ncon <= 0 || niter < 0)
            return METIS_ERROR_INPUT;
        if (!validate_csr(nvtxs, ncon, xadj, adjncy, vwgt) || !valid_part(nvtxs, part, 2) ||
            !totals_in_range(nvtxs, ncon, vwgt))
            return METIS_ERROR_INPUT;
        ContextGuard cx;
        hipStream_t s = cx->stream;
        DevGraph g = upload_graph(*cx.c, nvtxs, ncon, xadj, adjncy, vwgt, adjwgt, nullptr);
        int32_t* dp = upload(*cx.c, part, nvtxs);
        const size_t kc = 2 * (size_t)ncon;
        std::vector<long long> fm(kc), om(kc);
        thresholds_from_targets(2, ncon, tpwgts, ubvec, fm.data(), om.data());
        long long* dfm = upload(*cx.c, fm.data(), (int64_t)kc);
        fm2_refine_device(*cx.c, g, dfm, niter, dp);
        unsigned long long* acc = cx->arena.get<unsigned long long>(1);
        int32_t* out = cx->arena.get<int32_t>(1);
        edge_cut(g, dp, acc, out, pick_group(g.n, g.m), s);
        download(*cx.c, part, dp, nvtxs);
        int32_t c = 0;
        download(*cx.c, &c, out, 1);
        RK_HIP(hipStreamSynchronize(s));
        if (cut) *cut = c;
        return METIS_OK;
    });
}

int32_t rmetis_hip_kway_device(idx_t nvtxs, idx_t ncon, const idx_t* d_xadj, const idx_t* d_adjncy,
                               const idx_t* d_vwgt, const idx_t* d_adjwgt, idx_t nparts, const real_t* tpwgts,
                               const real_t* ubvec, const idx_t* options, idx_t* objval, idx_t* d_part,
                               rmetis_hip_stats* stats) {
    return guarded([&]() -> int32_t {
        if (!d_xadj || !d_adjncy || !d_part) return METIS_ERROR_INPUT;
        if (nvtxs <= 0 || ncon <= 0 || nparts < 2) return METIS_ERROR_INPUT;
        const Options o = parse_options(options);
        ContextGuard cx;
        DevGraph g;
        g.n = nvtxs;
        g.ncon = ncon;
        RK_HIP(hipMemcpyAsync(&g.m, d_xadj + nvtxs, sizeof(int32_t), hipMemcpyDeviceToHost, cx->stream));
        RK_HIP(hipStreamSynchronize(cx->stream));
        g.xadj = (int32_t*)d_xadj;
        g.adjncy = (int32_t*)d_adjncy;
        g.vwgt = (int32_t*)d_vwgt;
        g.adjwgt = (int32_t*)d_adjwgt;
        KwayStats ks;
        const int32_t obj = kway_device(*cx.c, g, nparts, tpwgts, ubvec, o, d_part, stats ? &ks : nullptr);
        if (stats) fill_stats(ks, stats);
        if (objval) *objval = obj;
        return METIS_OK;
    });
}

int32_t rmetis_hip_kway_device_noalloc(idx_t nvtxs, idx_t ncon, const idx_t* d_xadj, const idx_t* d_adjncy,
                                       const idx_t* d_vwgt, const idx_t* d_adjwgt, idx_t nparts,
                                       const real_t* tpwgts, const real_t* ubvec, const idx_t* options,
                                       idx_t* objval, idx_t* d_part) {
    // rmetis_hip_kway_device with every device buffer carved from `scratch`
    // (no hipMalloc, hipFree or pinned allocation; the null stream; host
    // staging in heap memory)
    return guarded([&]() -> int32_t {
        if (!d_xadj || !d_adjncy || !d_part) return METIS_ERROR_INPUT;
        if (nvtxs <= 0 || ncon <= 0 || nparts < 2) return METIS_ERROR_INPUT;
        const Options o = parse_options(options);
        int cnt = 0;
        if (hipGetDeviceCount(&cnt) != hipSuccess || cnt <= 0)
            throw Error(-4, "no HIP device is available (the k-way path runs only on the GPU)");
        int dev = 0;
        RK_HIP(hipGetDevice(&dev));
        if (dev < 0 || dev >= kMaxDevices) throw Error(-4, "device index out of range");
        std::lock_guard<std::mutex> lk(g_scratch_mu[dev]);
        void* ws = scratch_address(dev);
        DevGraph g;
        g.n = nvtxs;
        g.ncon = ncon;
        RK_HIP(hipMemcpy(&g.m, d_xadj + nvtxs, sizeof(int32_t), hipMemcpyDeviceToHost));
        if (g.m < 0) return METIS_ERROR_INPUT;
        const size_t need = kway_first_level_bytes(nvtxs, g.m, nparts, ncon);
        if (need > kScratchBytes)
            throw Error(-3, "graph needs at least " + std::to_string((need + (1 << 20) - 1) >> 20) +
                                " MiB of device workspace; the allocation-free entry point has " +
                                std::to_string(kScratchBytes >> 20) + " MiB");
        g.xadj = (int32_t*)d_xadj;
        g.adjncy = (int32_t*)d_adjncy;
        g.vwgt = (int32_t*)d_vwgt;
        g.adjwgt = (int32_t*)d_adjwgt;
        Context cx;
        cx.device = dev;
        cx.stream = nullptr;   // the null stream: no stream is created
        cx.pageable = true;
        cx.arena.attach(ws, kScratchBytes, true);
        int32_t obj = 0;
        try {
            obj = kway_device(cx, g, nparts, tpwgts, ubvec, o, d_part, nullptr);
        } catch (...) {
            // work already queued still uses the workspace: let it finish
            // before the next call may take the lock
            (void)hipStreamSynchronize(nullptr);
            throw;
        }
        if (objval) *objval = obj;
        return METIS_OK;
    });
}

int32_t rmetis_hip_kway_batch_device(idx_t ngraphs, const idx_t* gvtx, idx_t ncon, const idx_t* d_xadj,
                                     const idx_t* d_adjncy, const idx_t* d_vwgt, const idx_t* d_adjwgt,
                                     const idx_t* nparts, const real_t* ubvec, const idx_t* options,
                                     const uint64_t* seeds, idx_t* objval, idx_t* d_part, rmetis_hip_stats* stats) {
    return guarded([&]() -> int32_t {
        if (!gvtx || !d_xadj || !d_adjncy || !nparts || !objval || !d_part) return METIS_ERROR_INPUT;
        if (ngraphs <= 0 || ncon <= 0 || gvtx[ngraphs] <= 0) return METIS_ERROR_INPUT;
        const Options o = parse_options(options);
        ContextGuard cx;
        DevGraph g;
        g.n = gvtx[ngraphs];
        g.ncon = ncon;
        RK_HIP(hipMemcpyAsync(&g.m, d_xadj + g.n, sizeof(int32_t), hipMemcpyDeviceToHost, cx->stream));
        RK_HIP(hipStreamSynchronize(cx->stream));
        g.xadj = (int32_t*)d_xadj;
        g.adjncy = (int32_t*)d_adjncy;
        g.vwgt = (int32_t*)d_vwgt;
        g.adjwgt = (int32_t*)d_adjwgt;
        KwayStats ks;
        kway_batch_device(*cx.c, g, ngraphs, gvtx, nparts, ubvec, o, seeds, d_part, objval, stats ? &ks : nullptr);
        if (stats) fill_stats(ks, stats);
        return METIS_OK;
    });
}

int32_t rmetis_hip_kway_batch(idx_t ngraphs, const idx_t* gvtx, idx_t ncon, const idx_t* xadj, const idx_t* adjncy,
                              const idx_t* vwgt, const idx_t* adjwgt, const idx_t* nparts, const real_t* ubvec,
                              const idx_t* options, const uint64_t* seeds, idx_t* objval, idx_t* part) {
    return guarded([&]() -> int32_t {
        if (!gvtx || !xadj || !adjncy || !nparts || !objval || !part) return METIS_ERROR_INPUT;
        if (ngraphs <= 0 || ncon <= 0) return METIS_ERROR_INPUT;
        // graph/validate.rs:16-142 per graph: offsets first (xadj[0] == 0 and
        // monotone over the whole union, so that no row loop reads outside
        // adjncy[0..xadj[n])), then every row stays inside its graph
        if (gvtx[0] != 0 || xadj[0] != 0) return METIS_ERROR_INPUT;
        for (int32_t i = 0; i < ngraphs; ++i)
            if (gvtx[i + 1] <= gvtx[i]) return METIS_ERROR_INPUT;
        for (int32_t v = 0; v < gvtx[ngraphs]; ++v)
            if (xadj[v] > xadj[v + 1]) return METIS_ERROR_INPUT;
        for (int32_t i = 0; i < ngraphs; ++i) {
            const int32_t v0 = gvtx[i], v1 = gvtx[i + 1];
            if (v1 <= v0) return METIS_ERROR_INPUT;
            if (nparts[i] < 2 || nparts[i] > v1 - v0) return METIS_ERROR_INPUT;
            for (int32_t v = v0; v < v1; ++v) {
                for (int32_t j = xadj[v]; j < xadj[v + 1]; ++j)
                    if (adjncy[j] < v0 || adjncy[j] >= v1 || adjncy[j] == v) return METIS_ERROR_INPUT;
            }
        }
        const int32_t n = gvtx[ngraphs];
        std::vector<int32_t> cnt(n, 0);
        for (int64_t j = 0; j < xadj[n]; ++j) cnt[adjncy[j]]++;
        for (int32_t v = 0; v < n; ++v)
            if (cnt[v] != xadj[v + 1] - xadj[v]) return METIS_ERROR_INPUT;
        if (vwgt)
            for (int64_t i = 0; i < (int64_t)n * ncon; ++i)
                if (vwgt[i] <= 0) return METIS_ERROR_INPUT;
        const Options o = parse_options(options);
        ContextGuard cx;
        DevGraph g = upload_graph(*cx.c, n, ncon, xadj, adjncy, vwgt, adjwgt, nullptr);
        int32_t* dp = cx->arena.get<int32_t>(n);
        std::vector<int32_t> obj(ngraphs);
        kway_batch_device(*cx.c, g, ngraphs, gvtx, nparts, ubvec, o, seeds, dp, obj.data(), nullptr);
        download(*cx.c, part, dp, n);
        RK_HIP(hipStreamSynchronize(cx->stream));
        std::copy(obj.begin(), obj.end(), objval);
        return METIS_OK;
    });
}

rmetis_hip_server* rmetis_hip_server_create(int32_t slots, int64_t memory_budget, int64_t max_batch_vertices) {
    rmetis_hip_server* out = nullptr;
    const int32_t rc = guarded([&]() -> int32_t {
        if (slots < 1) return METIS_ERROR_INPUT;
        if (slots > kMaxServeSlots) throw Error(METIS_ERROR_INPUT, "a server has at most " +
                                                               std::to_string(kMaxServeSlots) + " slots");
        int cnt = 0;
        if (hipGetDeviceCount(&cnt) != hipSuccess || cnt <= 0)
            throw Error(-4, "no HIP device is available (the k-way path runs only on the GPU)");
        int dev = 0;
        RK_HIP(hipGetDevice(&dev));
        if (memory_budget <= 0) {
            size_t free_b = 0, total_b = 0;
            RK_HIP(hipMemGetInfo(&free_b, &total_b));
            memory_budget = (int64_t)(free_b / 10 * 8);
        }
        // large batches keep more graphs' admissions (one workgroup each) on
        // the device at once; the memory budget still bounds what runs
        if (max_batch_vertices <= 0) max_batch_vertices = 1ll << 25;
        auto* h = new rmetis_hip_server();
        h->device = dev;
        h->s = std::make_unique<Server>(dev, slots, memory_budget, max_batch_vertices);
        out = h;
        return METIS_OK;
    });
    return rc == METIS_OK ? out : nullptr;
}

int64_t rmetis_hip_server_submit(rmetis_hip_server* server, idx_t nvtxs, idx_t ncon, const idx_t* xadj,
                                 const idx_t* adjncy, const idx_t* vwgt, const idx_t* adjwgt, idx_t nparts,
                                 const real_t* ubvec, const idx_t* options, uint64_t seed) {
    int64_t id = 0;
    const int32_t rc = guarded([&]() -> int32_t {
        if (!server || !xadj || !adjncy) return METIS_ERROR_INPUT;
        if (nvtxs <= 0 || ncon <= 0 || nparts < 2) return METIS_ERROR_INPUT;           // c_api.rs:157-167
        if (!validate_csr(nvtxs, ncon, xadj, adjncy, vwgt)) return METIS_ERROR_INPUT;   // validate.rs:16-142
        if (nparts > nvtxs) return METIS_ERROR_INPUT;                                 // kway.rs:34-38
        const Options o = parse_options(options);
        if (!o.rtype_greedy)
            throw Error(-2, "the batched and served forms refine greedily: set options[METIS_OPTION_RTYPE]=1 "
                            "(Greedy) and options[METIS_OPTION_PTYPE]=0 so that the array is read; k-way FM runs "
                            "through RMETIS_PartGraphKway and rmetis_hip_kway_device");
        std::vector<int64_t> tot(ncon, 0);
        for (int64_t v = 0; v < nvtxs; ++v)
            for (int c = 0; c < ncon; ++c) tot[c] += vwgt ? vwgt[v * ncon + c] : 1;
        for (int c = 0; c < ncon; ++c)
            if (tot[c] > INT32_MAX)
                throw Error(-4, "total vertex weight of a constraint exceeds the int32 range");
        auto r = std::make_shared<ServeRequest>();
        r->n = nvtxs;
        r->ncon = ncon;
        r->k = nparts;
        const int64_t m = xadj[nvtxs];
        r->xadj.assign(xadj, xadj + nvtxs + 1);
        r->adjncy.assign(adjncy, adjncy + m);
        if (vwgt) r->vwgt.assign(vwgt, vwgt + (int64_t)nvtxs * ncon);
        if (adjwgt) r->adjwgt.assign(adjwgt, adjwgt + m);
        if (ubvec) r->ubvec.assign(ubvec, ubvec + ncon);
        r->opt = o;
        r->seed = seed != 0 ? seed : o.seed;
        id = server->s->submit(r);
        return METIS_OK;
    });
    return rc == METIS_OK ? id : (int64_t)rc;
}

int64_t rmetis_hip_server_submit_device(rmetis_hip_server* server, idx_t nvtxs, idx_t ncon, const idx_t* d_xadj,
                                        const idx_t* d_adjncy, const idx_t* d_vwgt, const idx_t* d_adjwgt,
                                        idx_t nparts, const real_t* ubvec, const idx_t* options, uint64_t seed,
                                        idx_t* d_part) {
    int64_t id = 0;
    const int32_t rc = guarded([&]() -> int32_t {
        if (!server || !d_xadj || !d_adjncy || !d_part) return METIS_ERROR_INPUT;
        if (nvtxs <= 0 || ncon <= 0 || nparts < 2 || nparts > nvtxs) return METIS_ERROR_INPUT;
        const Options o = parse_options(options);
        if (!o.rtype_greedy)
            throw Error(-2, "the batched and served forms refine greedily: set options[METIS_OPTION_RTYPE]=1 "
                            "(Greedy) and options[METIS_OPTION_PTYPE]=0 so that the array is read; k-way FM runs "
                            "through RMETIS_PartGraphKway and rmetis_hip_kway_device");
        int32_t m = 0;
        RK_HIP(hipMemcpy(&m, d_xadj + nvtxs, sizeof(int32_t), hipMemcpyDeviceToHost));
        auto r = std::make_shared<ServeRequest>();
        r->on_device = true;
        r->n = nvtxs;
        r->ncon = ncon;
        r->k = nparts;
        r->m = m;
        r->d_xadj = d_xadj;
        r->d_adjncy = d_adjncy;
        r->d_vwgt = d_vwgt;
        r->d_adjwgt = d_adjwgt;
        r->d_part = d_part;
        if (ubvec) r->ubvec.assign(ubvec, ubvec + ncon);
        r->opt = o;
        r->seed = seed != 0 ? seed : o.seed;
        id = server->s->submit(r);
        return METIS_OK;
    });
    return rc == METIS_OK ? id : (int64_t)rc;
}

int32_t rmetis_hip_server_result(rmetis_hip_server* server, int64_t id, int32_t wait, idx_t* part, idx_t* objval,
                                 double* times_ms, int32_t* batch_graphs) {
    return guarded([&]() -> int32_t {
        if (!server) return METIS_ERROR_INPUT;
        // wait: 0 = poll, 1 = until done, > 1 = at most that many milliseconds
        std::shared_ptr<ServeRequest> r = server->s->result(id, wait == 0 ? 0 : (wait == 1 ? -1 : (int64_t)wait));
        if (!r) return 0;
        if (r->status != METIS_OK) throw Error(r->status, r->error);
        if (part && !r->on_device) std::copy(r->part.begin(), r->part.end(), part);
        if (objval) *objval = r->objval;
        if (times_ms) {
            using ms = std::chrono::duration<double, std::milli>;
            times_ms[0] = ms(r->t_start - r->t_submit).count();
            times_ms[1] = ms(r->t_done - r->t_start).count();
            times_ms[2] = ms(r->t_done - r->t_submit).count();
        }
        if (batch_graphs) *batch_graphs = r->batch_graphs;
        return METIS_OK;
    });
}

int32_t rmetis_hip_server_stats(rmetis_hip_server* server, int64_t* out) {
    return guarded([&]() -> int32_t {
        if (!server || !out) return METIS_ERROR_INPUT;
        const ServerStats st = server->s->stats();
        out[0] = st.batches;
        out[1] = st.served;
        out[2] = st.inflight_batches;
        out[3] = st.max_inflight_batches;
        out[4] = st.max_batch_graphs;
        return METIS_OK;
    });
}

void rmetis_hip_server_destroy(rmetis_hip_server* server) {
    if (!server) return;
    (void)hipSetDevice(server->device);
    delete server;
}

int32_t rmetis_hip_edge_cut(idx_t nvtxs, const idx_t* xadj, const idx_t* adjncy, const idx_t* adjwgt,
                            const idx_t* part, idx_t* cut) {
    return guarded([&]() -> int32_t {
        if (!xadj || !adjncy || !part || !cut || nvtxs <= 0) return METIS_ERROR_INPUT;
        if (!validate_csr(nvtxs, 1, xadj, adjncy, nullptr)) return METIS_ERROR_INPUT;
        ContextGuard cx;
        DevGraph g = upload_graph(*cx.c, nvtxs, 1, xadj, adjncy, nullptr, adjwgt, nullptr);
        int32_t* dp = upload(*cx.c, part, nvtxs);
        auto* acc = cx->arena.get<unsigned long long>(1);
        int32_t* out = cx->arena.get<int32_t>(1);
        edge_cut(g, dp, acc, out, pick_group(g.n, g.m), cx->stream);
        download(*cx.c, cut, out, 1);
        RK_HIP(hipStreamSynchronize(cx->stream));
        return METIS_OK;
    });
}

int32_t rmetis_hip_partition_weights(idx_t nvtxs, idx_t ncon, const idx_t* vwgt, const idx_t* part,
                                     idx_t nparts, idx_t* pw) {
    return guarded([&]() -> int32_t {
        if (!part || !pw || nvtxs <= 0 || ncon <= 0 || nparts <= 0) return METIS_ERROR_INPUT;
        if (!valid_part(nvtxs, part, nparts)) return METIS_ERROR_INPUT;
        ContextGuard cx;
        DevGraph g;
        g.n = nvtxs;
        g.ncon = ncon;
        g.vwgt = upload(*cx.c, vwgt, (int64_t)nvtxs * ncon);
        int32_t* dp = upload(*cx.c, part, nvtxs);
        int32_t* dpw = cx->arena.get<int32_t>((int64_t)nparts * ncon);
        part_weights(g, dp, nparts, dpw, cx->stream);
        download(*cx.c, pw, dpw, (int64_t)nparts * ncon);
        RK_HIP(hipStreamSynchronize(cx->stream));
        return METIS_OK;
    });
}

int32_t rmetis_hip_gain_table(idx_t nvtxs, const idx_t* xadj, const idx_t* adjncy, const idx_t* adjwgt,
                              const idx_t* part, idx_t nparts, idx_t* rec, idx_t* g0, idx_t* cand_part,
                              idx_t* cand_gain) {
    return guarded([&]() -> int32_t {
        if (!xadj || !adjncy || !part || !rec || !g0 || !cand_part || !cand_gain || nvtxs <= 0 || nparts < 1)
            return METIS_ERROR_INPUT;
        if (!validate_csr(nvtxs, 1, xadj, adjncy, nullptr) || !valid_part(nvtxs, part, nparts))
            return METIS_ERROR_INPUT;
        ContextGuard cx;
        DevGraph g = upload_graph(*cx.c, nvtxs, 1, xadj, adjncy, nullptr, adjwgt, nullptr);
        int32_t* dp = upload(*cx.c, part, nvtxs);
        Buffers B = alloc_buffers(*cx.c, g.n, g.m, nparts, 1);
        int4* vrec = cx->arena.get<int4>(nvtxs);
        int32_t* dg0 = cx->arena.get<int32_t>(nvtxs);
        int32_t* gains = cx->arena.get<int32_t>(g.m);
        RK_HIP(hipMemsetAsync(dg0, 0, sizeof(int32_t) * nvtxs, cx->stream));
        int32_t* heavy = cx->arena.get<int32_t>(nvtxs);
        const int nheavy = heavy_list_sync(g, heavy, B.rb, cx->stream);
        gain_table(g, nparts, nheavy, heavy, B.rb, dp, vrec, dg0, gains, cx->stream);
        download(*cx.c, (int4*)rec, vrec, nvtxs);
        download(*cx.c, g0, dg0, nvtxs);
        download(*cx.c, cand_part, B.rb.pool, g.m);
        download(*cx.c, cand_gain, gains, g.m);
        RK_HIP(hipStreamSynchronize(cx->stream));
        return METIS_OK;
    });
}

int32_t rmetis_hip_greedy_pass(idx_t nvtxs, idx_t ncon, const idx_t* xadj, const idx_t* adjncy,
                               const idx_t* vwgt, const idx_t* adjwgt, idx_t* part, idx_t nparts,
                               const double* tpwgts, const double* ubvec, idx_t mode, idx_t* pw, idx_t* moves,
                               idx_t* nmoves) {
    return guarded([&]() -> int32_t {
        if (!xadj || !adjncy || !part || !tpwgts || !ubvec || !pw || nvtxs <= 0 || ncon <= 0 || nparts < 1 ||
            (mode != 0 && mode != 1))
            return METIS_ERROR_INPUT;
        if (!validate_csr(nvtxs, ncon, xadj, adjncy, vwgt) || !valid_part(nvtxs, part, nparts))
            return METIS_ERROR_INPUT;
        std::vector<int64_t> tot;
        if (!totals_in_range(nvtxs, ncon, vwgt, &tot)) return METIS_ERROR_INPUT;
        for (int64_t i = 0; i < (int64_t)nparts * ncon; ++i)
            if (pw[i] < 0 || pw[i] > tot[i % ncon]) return METIS_ERROR_INPUT;
        ContextGuard cx;
        hipStream_t s = cx->stream;
        DevGraph g = upload_graph(*cx.c, nvtxs, ncon, xadj, adjncy, vwgt, adjwgt, nullptr);
        int32_t* dp = upload(*cx.c, part, nvtxs);
        Buffers B = alloc_buffers(*cx.c, g.n, g.m, nparts, ncon);
        const size_t kc = (size_t)nparts * ncon;
        std::vector<long long> fm(kc), om(kc);
        thresholds_from_targets(nparts, ncon, tpwgts, ubvec, fm.data(), om.data());
        int32_t* heavy = cx->arena.get<int32_t>(nvtxs);
        const int nheavy = heavy_list_sync(g, heavy, B.rb, s);
        RK_HIP(hipMemcpyAsync(B.rb.fmax, fm.data(), kc * sizeof(long long), hipMemcpyHostToDevice, s));
        RK_HIP(hipMemcpyAsync(B.rb.omax, om.data(), kc * sizeof(long long), hipMemcpyHostToDevice, s));
        RK_HIP(hipMemcpyAsync(B.rb.pw, pw, kc * sizeof(int32_t), hipMemcpyHostToDevice, s));
        RK_HIP(hipMemsetAsync(B.rb.done, 0, sizeof(int32_t), s));
        RK_HIP(hipMemsetAsync(B.rb.passes, 0, sizeof(int32_t), s));
        refine_pass(g, nparts, mode, nheavy, heavy, B.rb, dp, s);
        int32_t nm = 0;
        download(*cx.c, &nm, B.rb.nmoves, 1);
        RK_HIP(hipStreamSynchronize(s));
        download(*cx.c, part, dp, nvtxs);
        download(*cx.c, pw, B.rb.pw, kc);
        if (moves && nm > 0) download(*cx.c, (int2*)moves, B.rb.moves, nm);
        RK_HIP(hipStreamSynchronize(s));
        if (nmoves) *nmoves = nm;
        return METIS_OK;
    });
}

int32_t rmetis_hip_greedy_refine(idx_t nvtxs, idx_t ncon, const idx_t* xadj, const idx_t* adjncy,
                                 const idx_t* vwgt, const idx_t* adjwgt, idx_t* part, idx_t nparts,
                                 const double* tpwgts, const double* ubvec, idx_t niter, idx_t* cut) {
    return guarded([&]() -> int32_t {
        if (!xadj || !adjncy || !part || !tpwgts || !ubvec || nvtxs <= 0 || ncon <= 0 || nparts < 1 || niter < 0)
            return METIS_ERROR_INPUT;
        if (!validate_csr(nvtxs, ncon, xadj, adjncy, vwgt) || !valid_part(nvtxs, part, nparts) ||
            !totals_in_range(nvtxs, ncon, vwgt))
            return METIS_ERROR_INPUT;
        ContextGuard cx;
        hipStream_t s = cx->stream;
        DevGraph g = upload_graph(*cx.c, nvtxs, ncon, xadj, adjncy, vwgt, adjwgt, nullptr);
        int32_t* dp = upload(*cx.c, part, nvtxs);
        Buffers B = alloc_buffers(*cx.c, g.n, g.m, nparts, ncon);
        const size_t kc = (size_t)nparts * ncon;
        std::vector<long long> fm(kc), om(kc);
        thresholds_from_targets(nparts, ncon, tpwgts, ubvec, fm.data(), om.data());
        int32_t* heavy = cx->arena.get<int32_t>(nvtxs);
        const int nheavy = heavy_list_sync(g, heavy, B.rb, s);
        RK_HIP(hipMemcpyAsync(B.rb.fmax, fm.data(), kc * sizeof(long long), hipMemcpyHostToDevice, s));
        RK_HIP(hipMemcpyAsync(B.rb.omax, om.data(), kc * sizeof(long long), hipMemcpyHostToDevice, s));
        refine_level(*cx.c, g, nparts, B.rb.fmax, B.rb.omax, niter, dp, nheavy, heavy, B, nullptr);
        edge_cut(g, dp, B.cut_acc, B.cut_out, pick_group(g.n, g.m), s);
        download(*cx.c, part, dp, nvtxs);
        int32_t c = 0;
        download(*cx.c, &c, B.cut_out, 1);
        RK_HIP(hipStreamSynchronize(s));
        if (cut) *cut = c;
        return METIS_OK;
    });
}

int32_t rmetis_hip_fm_refine(idx_t nvtxs, idx_t ncon, const idx_t* xadj, const idx_t* adjncy, const idx_t* vwgt,
                             const idx_t* adjwgt, idx_t* part, idx_t nparts, const double* tpwgts,
                             const double* ubvec, idx_t niter, idx_t* cut) {
    // fm_kway_refine, src/refine/fm_kway.rs:12-30
    return guarded([&]() -> int32_t {
        if (!xadj || !adjncy || !part || !tpwgts || !ubvec || nvtxs <= 0 || ncon <= 0 || nparts < 1 || niter < 0)
            return METIS_ERROR_INPUT;
        if (!validate_csr(nvtxs, ncon, xadj, adjncy, vwgt) || !valid_part(nvtxs, part, nparts) ||
            !totals_in_range(nvtxs, ncon, vwgt))
            return METIS_ERROR_INPUT;
        ContextGuard cx;
        hipStream_t s = cx->stream;
        DevGraph g = upload_graph(*cx.c, nvtxs, ncon, xadj, adjncy, vwgt, adjwgt, nullptr);
        int32_t* dp = upload(*cx.c, part, nvtxs);
        Buffers B = alloc_buffers(*cx.c, g.n, g.m, nparts, ncon);
        FmBufs F = alloc_fm_buffers(*cx.c, g.n, g.m, nparts, ncon);
        const size_t kc = (size_t)nparts * ncon;
        std::vector<long long> fm(kc), om(kc);
        thresholds_from_targets(nparts, ncon, tpwgts, ubvec, fm.data(), om.data());
        int32_t* heavy = cx->arena.get<int32_t>(nvtxs);
        const int nheavy = heavy_list_sync(g, heavy, B.rb, s);
        RK_HIP(hipMemcpyAsync(B.rb.fmax, fm.data(), kc * sizeof(long long), hipMemcpyHostToDevice, s));
        fm_refine_level(g, nparts, B.rb.fmax, niter, dp, nheavy, heavy, B.rb, F, s, false, false);
        edge_cut(g, dp, B.cut_acc, B.cut_out, pick_group(g.n, g.m), s);
        download(*cx.c, part, dp, nvtxs);
        int32_t c = 0, err = 0;
        download(*cx.c, &c, B.cut_out, 1);
        download(*cx.c, &err, F.err, 1);
        RK_HIP(hipStreamSynchronize(s));
        if (err) throw Error(-4, "FM pass exceeded its pop budget");
        if (cut) *cut = c;
        return METIS_OK;
    });
}

int32_t rmetis_hip_match(idx_t nvtxs, const idx_t* xadj, const idx_t* adjncy, const idx_t* adjwgt, idx_t ctype,
                         uint64_t seed, idx_t* matching) {
    return guarded([&]() -> int32_t {
        if (!xadj || !adjncy || !matching || nvtxs <= 0) return METIS_ERROR_INPUT;
        if (!validate_csr(nvtxs, 1, xadj, adjncy, nullptr)) return METIS_ERROR_INPUT;
        ContextGuard cx;
        DevGraph g = upload_graph(*cx.c, nvtxs, 1, xadj, adjncy, nullptr, adjwgt, nullptr);
        Buffers B = alloc_buffers(*cx.c, g.n, g.m, 2, 1);
        match_graph(g, ctype != 0, seed, B.cb, cx->stream);
        download(*cx.c, matching, B.cb.match, nvtxs);
        RK_HIP(hipStreamSynchronize(cx->stream));
        return METIS_OK;
    });
}

int32_t rmetis_hip_match_capped(idx_t nvtxs, idx_t ncon, const idx_t* xadj, const idx_t* adjncy, const idx_t* vwgt,
                                const idx_t* adjwgt, idx_t ctype, uint64_t seed, const int64_t* cap,
                                idx_t* matching) {
    return guarded([&]() -> int32_t {
        if (!xadj || !adjncy || !vwgt || !cap || !matching || nvtxs <= 0 || ncon <= 0) return METIS_ERROR_INPUT;
        if (!validate_csr(nvtxs, ncon, xadj, adjncy, vwgt)) return METIS_ERROR_INPUT;
        for (int32_t c = 0; c < ncon; ++c)
            if (cap[c] < 0 || cap[c] == LLONG_MAX) return METIS_ERROR_INPUT;
        ContextGuard cx;
        DevGraph g = upload_graph(*cx.c, nvtxs, ncon, xadj, adjncy, vwgt, adjwgt, nullptr);
        Buffers B = alloc_buffers(*cx.c, g.n, g.m, 2, ncon);
        long long* d_cap = cx->arena.get<long long>(ncon);
        long long* h = (long long*)cx->upload_stage(sizeof(long long) * ncon);
        for (int32_t c = 0; c < ncon; ++c) h[c] = cap[c];
        RK_HIP(hipMemcpyAsync(d_cap, h, sizeof(long long) * ncon, hipMemcpyHostToDevice, cx->stream));
        match_graph(g, ctype != 0, seed, B.cb, cx->stream, nullptr, d_cap);
        download(*cx.c, matching, B.cb.match, nvtxs);
        RK_HIP(hipStreamSynchronize(cx->stream));
        return METIS_OK;
    });
}

int32_t rmetis_hip_build_coarse_graph(idx_t nvtxs, idx_t ncon, const idx_t* xadj, const idx_t* adjncy,
                                      const idx_t* vwgt, const idx_t* adjwgt, const idx_t* vsize,
                                      const idx_t* matching, idx_t* cxadj, idx_t* cadjncy, idx_t* cadjwgt,
                                      idx_t* cvwgt, idx_t* cvsize, idx_t* cmap, idx_t* cn_out, idx_t* cm_out,
                                      idx_t* flags) {
    return guarded([&]() -> int32_t {
        if (!xadj || !adjncy || !matching || !cxadj || !cadjncy || !cadjwgt || !cmap || nvtxs <= 0 || ncon <= 0)
            return METIS_ERROR_INPUT;
        if (!validate_csr(nvtxs, ncon, xadj, adjncy, vwgt)) return METIS_ERROR_INPUT;
        for (int32_t v = 0; v < nvtxs; ++v) {
            const int32_t u = matching[v];
            if (u < 0 || u >= nvtxs || matching[u] != v) return METIS_ERROR_INPUT;  // symmetric only
        }
        ContextGuard cx;
        hipStream_t s = cx->stream;
        DevGraph g = upload_graph(*cx.c, nvtxs, ncon, xadj, adjncy, vwgt, adjwgt, vsize);
        Buffers B = alloc_buffers(*cx.c, g.n, g.m, 2, ncon);
        int32_t* dm = upload(*cx.c, matching, nvtxs);
        int32_t* dcmap = cx->arena.get<int32_t>(nvtxs);
        struct A : CoarseAlloc {
            Arena& a;
            explicit A(Arena& ar) : a(ar) {}
            int32_t* alloc_i32(int64_t c) override { return a.get<int32_t>(c); }
        } al(cx->arena);
        DevGraph cg;
        const ContractResult cr = contract(g, dm, dcmap, cg, true, B.cb, al, nullptr, s);
        const int32_t cn = cr.cn, cm = cr.cm;
        const bool not_one = cr.adjwgt_not_one;
        download(*cx.c, cxadj, cg.xadj, (int64_t)cn + 1);
        download(*cx.c, cadjncy, cg.adjncy, cm);
        download(*cx.c, cadjwgt, cg.adjwgt, cm);
        if (cg.vwgt && cvwgt) download(*cx.c, cvwgt, cg.vwgt, (int64_t)cn * ncon);
        if (cg.vsize && cvsize) download(*cx.c, cvsize, cg.vsize, cn);
        download(*cx.c, cmap, dcmap, nvtxs);
        RK_HIP(hipStreamSynchronize(s));
        if (cn_out) *cn_out = cn;
        if (cm_out) *cm_out = cm;
        if (flags)
            *flags = (cg.vwgt ? 1 : 0) | (cg.vsize ? 2 : 0) | ((adjwgt != nullptr || not_one) ? 4 : 0);
        return METIS_OK;
    });
}

int32_t rmetis_hip_validate_csr(idx_t nvtxs, idx_t ncon, const idx_t* xadj, int64_t xadj_len,
                                const idx_t* adjncy, int64_t adjncy_len, const idx_t* vwgt, int64_t vwgt_len,
                                const idx_t* adjwgt, int64_t adjwgt_len) {
    // graph/validate.rs:16-142 with explicit array lengths (the Rust Vec lengths).
    return guarded([&]() -> int32_t {
        if (nvtxs <= 0 || ncon <= 0 || !xadj) return METIS_ERROR_INPUT;
        if (xadj_len != (int64_t)nvtxs + 1) return METIS_ERROR_INPUT;
        if (xadj[0] != 0) return METIS_ERROR_INPUT;
        for (int32_t i = 0; i < nvtxs; ++i)
            if (xadj[i] > xadj[i + 1]) return METIS_ERROR_INPUT;
        if (adjncy_len != (int64_t)xadj[nvtxs]) return METIS_ERROR_INPUT;
        if (adjncy_len > 0 && !adjncy) return METIS_ERROR_INPUT;
        if (adjwgt && adjwgt_len != adjncy_len) return METIS_ERROR_INPUT;
        if (vwgt && vwgt_len != (int64_t)nvtxs * ncon) return METIS_ERROR_INPUT;
        return validate_csr(nvtxs, ncon, xadj, adjncy, vwgt) ? METIS_OK : METIS_ERROR_INPUT;
    });
}

int32_t rmetis_hip_parse_options(const idx_t* options, int64_t* out) {
    // c_api.rs:211-222 + types.rs:190-224; out[13] = ctype, iptype, rtype,
    // objtype, no2hop, ncuts, niter, seed, minconn, contig, ufactor, numbering, dbglvl
    if (!out) return METIS_ERROR_INPUT;
    const Options o = parse_options(options);
    const int64_t v[13] = {o.ctype_shem, o.iptype,  o.rtype_greedy, o.objtype,  o.no2hop,
                           o.ncuts,      o.niter,   (int64_t)o.seed, o.minconn, o.contig,
                           o.ufactor,    o.numbering, o.dbglvl};
    for (int i = 0; i < 13; ++i) out[i] = v[i];
    return METIS_OK;
}

int32_t rmetis_hip_grow_partition(idx_t nvtxs, idx_t ncon, const idx_t* xadj, const idx_t* adjncy,
                                  const idx_t* vwgt, idx_t nparts, const double* tpwgts, uint64_t seed,
                                  idx_t* part) {
    // initial/grow.rs:12-103 with SmallRng::seed_from_u64(seed) (kway.rs:62)
    return guarded([&]() -> int32_t {
        if (!xadj || !adjncy || !tpwgts || !part || nvtxs <= 0 || ncon <= 0 || nparts < 1) return METIS_ERROR_INPUT;
        if (!validate_csr(nvtxs, ncon, xadj, adjncy, vwgt)) return METIS_ERROR_INPUT;
        if (nparts > nvtxs) return METIS_ERROR_INPUT;   // the reference's seed loop would not end
        Rng rng(seed);
        ContextGuard cx;
        DevGraph g = upload_graph(*cx.c, nvtxs, ncon, xadj, adjncy, vwgt, nullptr, nullptr);
        int32_t* d_part = cx->arena.get<int32_t>(nvtxs);
        GrowJob job;
        job.g = g;
        job.n = nvtxs;
        job.m = xadj[nvtxs];
        job.k = nparts;
        job.tpw.assign(tpwgts, tpwgts + (size_t)nparts * ncon);
        job.seeds = grow_seeds(nvtxs, nparts, rng);
        job.out = d_part;
        {
            // rows of a user graph may repeat an entry (the validation allows
            // it): one pass with a per-vertex marker of the current row
            std::vector<int32_t> mark(nvtxs, -1);
            bool rep = false;
            for (int32_t v = 0; v < nvtxs && !rep; ++v)
                for (int32_t j = xadj[v]; j < xadj[v + 1]; ++j) {
                    if (mark[adjncy[j]] == v) {
                        rep = true;
                        break;
                    }
                    mark[adjncy[j]] = v;
                }
            job.norep = !rep;
        }
        grow_on_device(*cx.c, std::vector<GrowJob>{std::move(job)});
        download(*cx.c, part, d_part, nvtxs);
        RK_HIP(hipStreamSynchronize(cx->stream));
        return METIS_OK;
    });
}

int32_t rmetis_hip_level_thresholds(idx_t nparts, idx_t ncon, const real_t* tpwgts, const idx_t* total,
                                    const real_t* ubvec, idx_t ufactor, int64_t* fmax, int64_t* omax) {
    // graph/mod.rs:392-410 targets, kway.rs:44-49 imbalance, greedy.rs:70,87,144 compares
    return guarded([&]() -> int32_t {
        if (!total || !fmax || !omax || nparts < 1 || ncon < 1) return METIS_ERROR_INPUT;
        std::vector<double> ub(ncon);
        for (int32_t c = 0; c < ncon; ++c) ub[c] = ubvec ? (double)ubvec[c] : 1.0 + (double)ufactor / 1000.0;
        std::vector<long long> f((size_t)nparts * ncon), o((size_t)nparts * ncon);
        level_thresholds(nparts, ncon, tpwgts, total, ub.data(), f.data(), o.data());
        for (size_t i = 0; i < f.size(); ++i) {
            fmax[i] = f[i];
            omax[i] = o[i];
        }
        return METIS_OK;
    });
}

int32_t rmetis_hip_get_tuning(int32_t* values, int32_t count) {
    if (!values || count <= 0) return 0;
    const int c = count < kTuneCount ? count : kTuneCount;
    for (int i = 0; i < c; ++i) values[i] = tune(i);
    return c;
}

int32_t rmetis_hip_set_tuning(const int32_t* values, int32_t count) {
    const char* why = "";
    if (!set_tuning(values, count, &why)) {
        g_last_error = why;
        return METIS_ERROR_INPUT;
    }
    return METIS_OK;
}

const char* rmetis_hip_tuning_source(void) { return tuning_source(); }

void rmetis_hip_debug_flags(int32_t flags) { set_debug_flags(flags); }

void* rmetis_hip_device_alloc(int64_t bytes) {
    void* p = nullptr;
    if (bytes <= 0 || hipMalloc(&p, (size_t)bytes) != hipSuccess) return nullptr;
    return p;
}

void rmetis_hip_device_free(void* p) {
    if (p) (void)hipFree(p);
}

int32_t rmetis_hip_memcpy(void* dst, const void* src, int64_t bytes, int32_t kind) {
    // kind: 1 host->device, 2 device->host, 3 device->device
    if (bytes < 0 || (bytes > 0 && (!dst || !src))) return METIS_ERROR_INPUT;
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                                     : kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    if (bytes > 0 && hipMemcpy(dst, src, (size_t)bytes, k) != hipSuccess) return METIS_ERROR;
    return METIS_OK;
}

int32_t rmetis_hip_device_synchronize(void) {
    return hipDeviceSynchronize() == hipSuccess ? METIS_OK : METIS_ERROR;
}

void rmetis_hip_profile_enable(int32_t on) { Profiler::get().enable(on != 0); }

void rmetis_hip_profile_filter(int32_t tag, int64_t size) { Profiler::get().filter(tag, (long long)size); }

int32_t rmetis_hip_profile_read(int32_t tag, int64_t size, int64_t* launches, double* ms, double* bytes) {
    if (!launches || !ms || !bytes || tag < 0 || tag >= kProfTags) return METIS_ERROR_INPUT;
    long long l = 0;
    Profiler::get().read(tag, size, &l, ms, bytes);
    *launches = l;
    return METIS_OK;
}

const char* rmetis_hip_last_error(void) { return g_last_error.c_str(); }

int32_t rmetis_hip_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

}  // extern "C"
